// Learned preconditioner of the device Newton-Krylov solver: a message-passing network on
// the bond graph, evaluated with BF16 MFMA (FP32 accumulation), trained on the GPU by its
// own backward kernels and Adam.  No counterpart in the reference, which calls
// IterativeSolvers.gmres! without a preconditioner (src/time_solvers/newton_krylov.jl:
// 448-450); it plugs into the flexible GMRES of pdx_nk.hip as a right preconditioner, with
// the point-block Jacobi solve as fallback.
//
// Network, per point i (internal Morton order), theta in float32 (master copy):
//   F_i (16)  = [D^-1 y / s, y / rms(y), constrained mask, normalised X, near-degree
//               ratio, bond-count ratio, 0, 0]      (s = rms of D^-1 y over free dofs)
//   H_i       = relu(W_e^T F_i + b_e)                                     16 -> 64
//   3 layers:   A_i = mean_{j in N(i)} H_j   (near bonds: L <= 0.6 delta, symmetric)
//               P_i, Q_i, G = means of H over the 64-point and 4096-point blocks of i and
//                             over the body
//               Z_i = W_l^T [H_i A_i P_i Q_i G] + b_l                     320 -> 64 (MFMA)
//               H_i = H_i + relu(Z_i)
//   O_i       = W_d^T H_i + b_d                                           64 -> 3
//   z         = D^-1 y + c s O  (free dofs),  z = y (constrained dofs)
// H and the concatenated layer inputs are stored in bfloat16; every sum is float32.
// Training pairs (y, x) with J x = y come from the plain solves: the Newton corrections
// (-r, du) and the Krylov directions (J z_j, z_j).  Loss per pair:
//   sum_free (O - t)^2 / sum_free t^2,   t = (x - D^-1 y) / (c s).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "pdx_nk.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define GF 64          // feature width
#define GFIN 16        // input features
#define GK (5 * GF)    // concatenated layer input
#define GL 3           // message-passing layers
#define GB1 64         // first pooling block
#define GB2 4096       // second pooling block
#define GPAD 512       // row padding of every activation buffer
#define GDW_CHUNK 512  // points per workgroup of the weight-gradient GEMM
#define G_NEAR_FRAC 0.6

// flat parameter layout (float32), identical to the CPU oracle's
#define OFF_WE 0
#define OFF_BE (OFF_WE + GFIN * GF)
#define OFF_W(l) (OFF_BE + GF + (l) * (GK * GF + GF))
#define OFF_B(l) (OFF_W(l) + GK * GF)
#define OFF_WD (OFF_BE + GF + GL * (GK * GF + GF))
#define OFF_BD (OFF_WD + GF * 3)
#define G_NPARAM (OFF_BD + 3)

// held-out loss below which the trained network is applied (1 = it explains nothing)
#define PDX_GNN_GATE 0.85

struct pdx_gnn {
    int64_t n = 0, npad = 0, nb1 = 0, nb2 = 0;
    // graph and static features
    bool ready = false;
    int64_t n_near = 0;
    int64_t *near_ptr = nullptr;
    int32_t *near = nullptr;
    float *stat = nullptr;  // 5 per point
    // parameters
    bool trained = false;
    double c_out = 1.0;
    // validation: the last tenth of the Krylov pairs is held out of training; the network is
    // applied only if it reduces the held-out error of the two-level preconditioner enough
    int64_t n_k_val = 0;
    double val_loss = -1.0;
    bool use_net = true;
    float *theta = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
    uint16_t *wT[GL] = {}, *wN[GL] = {};  // bf16: out-major [64][320], in-major [320][64]
    int64_t adam_t = 0;
    // activations (npad rows)
    float *F = nullptr, *Z0 = nullptr;
    uint16_t *H[GL + 1] = {};
    uint16_t *X[GL] = {};
    float *Z[GL] = {};
    float *O = nullptr;
    double *q = nullptr;
    float *S1 = nullptr, *S2 = nullptr, *Sg = nullptr;  // pooled sums, 192 columns
    double *scal = nullptr;                               // device scalars
    double *dpart = nullptr;                              // double partials
    // backward
    float *dH = nullptr, *dH2 = nullptr, *dZf = nullptr, *dO = nullptr;
    uint16_t *dZ = nullptr;
    float *dX = nullptr;
    float *wpart = nullptr;  // weight-gradient partials
    int64_t wpart_len = 0;
    // training pairs (float32, internal order)
    bool collecting = false;
    int64_t cap_k = 0, n_k = 0, seen_k = 0;
    float *Ky = nullptr, *Kx = nullptr;
    int64_t cap_n = 0, n_n = 0;
    float *Ny = nullptr, *Nx = nullptr;
    double *ys = nullptr, *xs = nullptr;  // FP64 sample
    std::mt19937_64 rng{12345};
};

// device scalars (scal[]): 0 s, 1 rms(y), 2 n_free, 3 sum r^2, 4 sum t^2, 5 loss accum,
// 6 dO scale, 7 c_out, 8 |x - q|^2 (output-scale estimate)
enum { SC_S = 0, SC_RY, SC_NFREE, SC_R2, SC_T2, SC_LOSS, SC_DSCALE, SC_C, SC_XQ, SC_N };

// ----------------------------------------------------------------------------------------
// helpers
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ bf16x8 load_bf8(const uint16_t *p)
{
    const u32x4 v = *reinterpret_cast<const u32x4 *>(p);
    return __builtin_bit_cast(bf16x8, v);
}

static inline dim3 g1(int64_t m, int bs = 256) { return dim3((unsigned)((m + bs - 1) / bs)); }

// ----------------------------------------------------------------------------------------
// graph
// ----------------------------------------------------------------------------------------
__global__ void k_near_count(int64_t n, const int64_t *slice_ptr, const int32_t *cnt,
                             const double *L, double lmax, int32_t *out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t b0 = slice_ptr[p >> 6] + (p & 63);
    int k = 0;
    for (int t = 0; t < cnt[p]; t++) k += (L[b0 + 64 * (int64_t)t] <= lmax) ? 1 : 0;
    out[p] = k;
}

__global__ void k_near_fill(int64_t n, const int64_t *slice_ptr, const int32_t *cnt,
                            const uint32_t *bw, const double *L, double lmax,
                            const int64_t *near_ptr, int32_t *near)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t b0 = slice_ptr[p >> 6] + (p & 63);
    int64_t o = near_ptr[p];
    for (int t = 0; t < cnt[p]; t++) {
        const int64_t b = b0 + 64 * (int64_t)t;
        if (L[b] <= lmax) near[o++] = (int32_t)(bw[b] & PDX_BW_NBR);
    }
}

// ----------------------------------------------------------------------------------------
// forward
// ----------------------------------------------------------------------------------------
// q = D^-1 y (point blocks, column-major 3x3), partial sums of q^2, y^2 and the free count
__global__ __launch_bounds__(256) void k_gnn_q(int64_t n, const double *dinv, const double *y,
                                               const uint8_t *cmask, double *q, double *part)
{
    __shared__ double sh[3][256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double sq = 0, sy = 0, nf = 0;
    if (i < n) {
        const double *R = dinv + 9 * i;
        const double a = y[3 * i], b = y[3 * i + 1], cc = y[3 * i + 2];
        const double qq[3] = {R[0] * a + R[3] * b + R[6] * cc, R[1] * a + R[4] * b + R[7] * cc,
                              R[2] * a + R[5] * b + R[8] * cc};
        for (int d = 0; d < 3; d++) {
            q[3 * i + d] = qq[d];
            if (!cmask[3 * i + d]) {
                sq += qq[d] * qq[d];
                sy += y[3 * i + d] * y[3 * i + d];
                nf += 1.0;
            }
        }
    }
    sh[0][threadIdx.x] = sq;
    sh[1][threadIdx.x] = sy;
    sh[2][threadIdx.x] = nf;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_gnn_q_final(int nblk, const double *part, double *scal)
{
    __shared__ double sh[3][256];
    double s[3] = {0, 0, 0};
    for (int b = threadIdx.x; b < nblk; b += 256)
        for (int k = 0; k < 3; k++) s[k] += part[3 * b + k];
    for (int k = 0; k < 3; k++) sh[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double nf = fmax(sh[2][0], 1.0);
        scal[SC_S] = sqrt(sh[0][0] / nf) + 1e-300;
        scal[SC_RY] = sqrt(sh[1][0] / nf) + 1e-300;
        scal[SC_NFREE] = sh[2][0];
    }
}

// encoder: F_i and H0_i = relu(W_e^T F_i + b_e); one wavefront per point, lane = feature
// (the per-point inputs are broadcast loads, the stores of H0 are coalesced)
__global__ __launch_bounds__(256) void k_gnn_encode(int64_t n, int64_t npad, const double *q,
                                                    const double *y, const uint8_t *cmask,
                                                    const float *stat, const double *scal,
                                                    const float *theta, float *Fo, float *Z0,
                                                    uint16_t *H0)
{
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int f = threadIdx.x & 63;
    if (p >= npad) return;
    if (p >= n) {
        H0[p * GF + f] = 0;
        if (Fo && f < GFIN) Fo[p * GFIN + f] = 0.f;
        if (Z0) Z0[p * GF + f] = 0.f;
        return;
    }
    const float is = (float)(1.0 / scal[SC_S]), iy = (float)(1.0 / scal[SC_RY]);
    float Fi[GFIN];
    for (int d = 0; d < 3; d++) {
        const bool c = cmask[3 * p + d] != 0;
        Fi[d] = c ? 0.f : (float)q[3 * p + d] * is;
        Fi[3 + d] = c ? 0.f : (float)y[3 * p + d] * iy;
        Fi[6 + d] = c ? 1.f : 0.f;
    }
    for (int k = 0; k < 5; k++) Fi[9 + k] = stat[5 * p + k];
    Fi[14] = Fi[15] = 0.f;
    if (Fo && f < GFIN) Fo[p * GFIN + f] = Fi[f];
    float z = theta[OFF_BE + f];
#pragma unroll
    for (int k = 0; k < GFIN; k++) z += Fi[k] * theta[OFF_WE + k * GF + f];
    if (Z0) Z0[p * GF + f] = z;
    H0[p * GF + f] = f2bf(fmaxf(z, 0.f));
}

// sums over 64-point blocks of `ng` column groups of 64 (source bf16 or float, row stride
// `ld`, first column `c0`); a workgroup of 4 x 64 ng threads per block: thread (g, col)
// sums rows g, g + 4, ..., the four partial sums meet in LDS
template <typename T>
__global__ __launch_bounds__(768) void k_gnn_pool1(int64_t n, int ng, const T *src, int ld, int c0,
                                                   float *S1)
{
    __shared__ float sh[4][3 * GF];
    const int64_t b = blockIdx.x;
    const int ncol = GF * ng;
    const int col = threadIdx.x % ncol, g = threadIdx.x / ncol;
    const int64_t p0 = b * GB1, p1 = min(n, p0 + GB1);
    float s = 0.f;
#pragma unroll 4
    for (int64_t p = p0 + g; p < p1; p += 4) {
        if constexpr (sizeof(T) == 2) s += bf2f(src[p * ld + c0 + col]);
        else s += src[p * ld + c0 + col];
    }
    sh[g][col] = s;
    __syncthreads();
    if (g == 0) S1[b * (GF * 3) + col] = (sh[0][col] + sh[1][col]) + (sh[2][col] + sh[3][col]);
}

// sums of 64 consecutive first-level sums (4096-point blocks)
__global__ void k_gnn_pool2(int64_t nb1, int ng, const float *S1, float *S2)
{
    const int64_t B = blockIdx.x;
    const int col = threadIdx.x;
    const int64_t b0 = B * (GB2 / GB1), b1 = min(nb1, b0 + GB2 / GB1);
    float s = 0.f;
    for (int64_t b = b0; b < b1; b++) s += S1[b * (GF * 3) + col];
    S2[B * (GF * 3) + col] = s;
}

__global__ void k_gnn_poolg(int64_t nb2, const float *S2, float *Sg)
{
    const int col = threadIdx.x;
    float s = 0.f;
    for (int64_t B = 0; B < nb2; B++) s += S2[B * (GF * 3) + col];
    Sg[col] = s;
}

__device__ __forceinline__ float blk_count(int64_t n, int64_t b, int64_t size)
{
    return (float)(min(n, (b + 1) * size) - b * size);
}

// X_i = [H_i, A_i, P_i, Q_i, G] in bf16; one wavefront (64 lanes = features) per point
__global__ __launch_bounds__(256) void k_gnn_concat(int64_t n, int64_t npad, const uint16_t *H,
                                                    const int64_t *near_ptr, const int32_t *near,
                                                    const float *S1, const float *S2,
                                                    const float *Sg, uint16_t *X)
{
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int f = threadIdx.x & 63;
    if (p >= npad) return;
    uint16_t *x = X + p * GK;
    if (p >= n) {
        for (int g = 0; g < 5; g++) x[g * GF + f] = 0;
        return;
    }
    const int64_t j0 = near_ptr[p], j1 = near_ptr[p + 1];
    float a = 0.f;
    // eight neighbours in flight: indices first, then their rows
    int64_t j = j0;
    for (; j + 8 <= j1; j += 8) {
        int32_t q[8];
#pragma unroll
        for (int u = 0; u < 8; u++) q[u] = near[j + u];
        float h[8];
#pragma unroll
        for (int u = 0; u < 8; u++) h[u] = bf2f(H[(int64_t)q[u] * GF + f]);
#pragma unroll
        for (int u = 0; u < 8; u++) a += h[u];
    }
    for (; j < j1; j++) a += bf2f(H[(int64_t)near[j] * GF + f]);
    if (j1 > j0) a /= (float)(j1 - j0);
    const int64_t b1 = p / GB1, b2 = p / GB2;
    x[f] = H[p * GF + f];
    x[GF + f] = f2bf(a);
    x[2 * GF + f] = f2bf(S1[b1 * (GF * 3) + f] / blk_count(n, b1, GB1));
    x[3 * GF + f] = f2bf(S2[b2 * (GF * 3) + f] / blk_count(n, b2, GB2));
    x[4 * GF + f] = f2bf(Sg[f] / (float)n);
}

// Z = X W + b, H_out = bf16(H_in + relu(Z)); a workgroup of four wavefronts computes 64
// rows (one 64-point pooling block), each wavefront 16 rows x 64 columns with
// v_mfma_f32_16x16x32_bf16 (lane l: A row l%16, B column l%16, k block 8*(l/16); D rows
// 4*(l/16)+i, column l%16).  Rows: npad (a multiple of 256).  With S1 given, the block's
// column sums of the new H (rows < n) are written for the next layer's pooling.
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_gnn_gemm_fwd(int64_t n, const uint16_t *X, const uint16_t *WT,
                                                      const float *bias, const uint16_t *Hin,
                                                      uint16_t *Hout, float *Z, float *S1)
{
    __shared__ float bsum[4][GF];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t blk0 = (int64_t)blockIdx.x * 64;
    const int64_t row0 = blk0 + 16 * wv;
    const int lr = lane & 15, lk = (lane >> 4) * 8;
    f32x4 acc[4];
#pragma unroll
    for (int c = 0; c < 4; c++) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int k0 = 0; k0 < GK; k0 += 32) {
        const bf16x8 a = load_bf8(X + (row0 + lr) * GK + k0 + lk);
        bf16x8 b[4];
#pragma unroll
        for (int c = 0; c < 4; c++) b[c] = load_bf8(WT + (int64_t)(16 * c + lr) * GK + k0 + lk);
#pragma unroll
        for (int c = 0; c < 4; c++) acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[c], acc[c], 0, 0, 0);
    }
    float bs[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t row = row0 + 4 * (lane >> 4) + i;
            const int col = 16 * c + lr;
            const float z = acc[c][i] + bias[col];
            if (TRAIN) Z[row * GF + col] = z;
            const uint16_t h = f2bf(bf2f(Hin[row * GF + col]) + fmaxf(z, 0.f));
            Hout[row * GF + col] = h;
            if (row < n) bs[c] += bf2f(h);
        }
    if (S1) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            bs[c] += __shfl_xor(bs[c], 16);
            bs[c] += __shfl_xor(bs[c], 32);
        }
        if (lane < 16)
#pragma unroll
            for (int c = 0; c < 4; c++) bsum[wv][16 * c + lane] = bs[c];
        __syncthreads();
        if (threadIdx.x < GF && blk0 < n)
            S1[(blk0 / GB1) * (GF * 3) + threadIdx.x] =
                (bsum[0][threadIdx.x] + bsum[1][threadIdx.x]) + (bsum[2][threadIdx.x] + bsum[3][threadIdx.x]);
    }
}

// decoder O = W_d^T H + b_d.  Apply: z = q + c s O (free), y (constrained).
// Train: O kept, partial sums of (O - t)^2 and t^2 with t = (x - q) / (c s).
// 16 lanes per point (4 features each, 8-byte loads), 16 points per workgroup.
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_gnn_decode(int64_t n, const uint16_t *H, const float *theta,
                                                    const double *q, const double *y,
                                                    const double *x, const uint8_t *cmask,
                                                    const double *scal, double *z, float *Oo,
                                                    double *part)
{
    __shared__ double sh[2][16];
    const int lane16 = threadIdx.x & 15, pl = threadIdx.x >> 4;
    const int64_t p = (int64_t)blockIdx.x * 16 + pl;
    double r2 = 0.0, t2 = 0.0;
    float O[3] = {0.f, 0.f, 0.f};
    if (p < n) {
        const uint2 hv = *reinterpret_cast<const uint2 *>(H + p * GF + 4 * lane16);
        const uint16_t hb[4] = {(uint16_t)(hv.x & 0xFFFF), (uint16_t)(hv.x >> 16),
                                (uint16_t)(hv.y & 0xFFFF), (uint16_t)(hv.y >> 16)};
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int f = 4 * lane16 + e;
            const float h = bf2f(hb[e]);
            O[0] += h * theta[OFF_WD + 3 * f];
            O[1] += h * theta[OFF_WD + 3 * f + 1];
            O[2] += h * theta[OFF_WD + 3 * f + 2];
        }
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1)
#pragma unroll
        for (int d = 0; d < 3; d++) O[d] += __shfl_xor(O[d], o);
    if (p < n && lane16 == 0) {
        const double cs = scal[SC_C] * scal[SC_S];
        for (int d = 0; d < 3; d++) {
            const int64_t k = 3 * p + d;
            const float od = O[d] + theta[OFF_BD + d];
            const bool c = cmask[k] != 0;
            if (!TRAIN) {
                z[k] = c ? y[k] : q[k] + cs * (double)od;
            } else {
                Oo[k] = od;
                if (!c) {
                    const double t = (x[k] - q[k]) / cs;
                    const double r = (double)od - t;
                    r2 += r * r;
                    t2 += t * t;
                }
            }
        }
    }
    if (TRAIN) {
        if (lane16 == 0) {
            sh[0][pl] = r2;
            sh[1][pl] = t2;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0, b = 0.0;
            for (int k = 0; k < 16; k++) {
                a += sh[0][k];
                b += sh[1][k];
            }
            part[2 * blockIdx.x] = a;
            part[2 * blockIdx.x + 1] = b;
        }
    }
}

// loss of one sample, the dO scale 2 / (B sum t^2)
__global__ void k_gnn_loss_final(int nblk, const double *part, double *scal, double inv_batch)
{
    __shared__ double sh[2][256];
    double a = 0, b = 0;
    for (int k = threadIdx.x; k < nblk; k += 256) {
        a += part[2 * k];
        b += part[2 * k + 1];
    }
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double t2 = fmax(sh[1][0], 1e-300);
        scal[SC_R2] = sh[0][0];
        scal[SC_T2] = t2;
        scal[SC_LOSS] += sh[0][0] / t2 * inv_batch;
        scal[SC_DSCALE] = 2.0 * inv_batch / t2;
    }
}

// ----------------------------------------------------------------------------------------
// backward
// ----------------------------------------------------------------------------------------
// dO = scale (O - t) on free dofs; dH = W_d dO.  One wavefront per point, lane = feature.
__global__ __launch_bounds__(256) void k_gnn_decode_bwd(int64_t n, int64_t npad, const float *theta,
                                                        const float *O, const double *q,
                                                        const double *x, const uint8_t *cmask,
                                                        const double *scal, float *dO, float *dH)
{
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int f = threadIdx.x & 63;
    if (p >= npad) return;
    float g[3] = {0.f, 0.f, 0.f};
    if (p < n) {
        const double cs = scal[SC_C] * scal[SC_S];
        for (int d = 0; d < 3; d++) {
            const int64_t k = 3 * p + d;
            if (!cmask[k]) {
                const double t = (x[k] - q[k]) / cs;
                g[d] = (float)(scal[SC_DSCALE] * ((double)O[k] - t));
            }
        }
    }
    if (f < 3) dO[3 * p + f] = g[f];
    dH[p * GF + f] = theta[OFF_WD + 3 * f] * g[0] + theta[OFF_WD + 3 * f + 1] * g[1] +
                     theta[OFF_WD + 3 * f + 2] * g[2];
}

// partial outer products sum_p A_p B_p^T and sum_p B_p over chunks of points: output
// [(ka + 1) x kb] per workgroup (row ka = bias gradient)
template <typename TA>
__global__ __launch_bounds__(256) void k_gnn_outer(int64_t npad, int64_t chunk, const TA *A, int ka,
                                                   const float *B, int kb, float *part)
{
    const int64_t p0 = (int64_t)blockIdx.x * chunk, p1 = min(npad, p0 + chunk);
    const int tot = (ka + 1) * kb;
    for (int e = threadIdx.x; e < tot; e += 256) {
        const int a = e / kb, b = e - a * kb;
        float s = 0.f;
        for (int64_t p = p0; p < p1; p++) {
            float av = 1.f;
            if (a < ka) {
                if constexpr (sizeof(TA) == 2) av = bf2f(A[p * ka + a]);
                else av = A[p * ka + a];
            }
            s += av * B[p * kb + b];
        }
        part[(int64_t)blockIdx.x * tot + e] = s;
    }
}

// out[e] += sum_w part[w * len + e]
__global__ void k_gnn_reduce_add(int64_t nw, int64_t len, const float *part, float *out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len) return;
    float s = 0.f;
    for (int64_t w = 0; w < nw; w++) s += part[w * len + e];
    out[e] += s;
}

// dZ = dH * [Z > 0] (bf16, and float when dZf is given); bias partials per 256 points.
// Workgroup of 256 threads = 4 row groups x 64 features over 256 rows.
__global__ __launch_bounds__(256) void k_gnn_relu_bwd(int64_t n, const float *dH, const float *Z,
                                                      uint16_t *dZ, float *dZf, float *bpart)
{
    __shared__ float sh[4][GF];
    const int f = threadIdx.x & 63, gq = threadIdx.x >> 6;
    const int64_t p0 = (int64_t)blockIdx.x * 256;
    float s = 0.f;
#pragma unroll 4
    for (int64_t p = p0 + gq; p < p0 + 256; p += 4) {
        const float g = (p < n && Z[p * GF + f] > 0.f) ? dH[p * GF + f] : 0.f;
        if (dZ) dZ[p * GF + f] = f2bf(g);
        if (dZf) dZf[p * GF + f] = g;
        s += g;
    }
    sh[gq][f] = s;
    __syncthreads();
    if (bpart && gq == 0) bpart[(int64_t)blockIdx.x * GF + f] = (sh[0][f] + sh[1][f]) + (sh[2][f] + sh[3][f]);
}

// dX = dZ W^T: [npad x 64] . [64 x 320]; one wavefront per 16 rows, all 320 columns
__global__ __launch_bounds__(256) void k_gnn_gemm_dx(const uint16_t *dZ, const uint16_t *WN,
                                                     float *dX)
{
    const int lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    const int lr = lane & 15, lk = (lane >> 4) * 8;
    f32x4 acc[GK / 16];
#pragma unroll
    for (int c = 0; c < GK / 16; c++) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < GF; k0 += 32) {
        const bf16x8 a = load_bf8(dZ + (row0 + lr) * GF + k0 + lk);
#pragma unroll
        for (int c = 0; c < GK / 16; c++) {
            const bf16x8 b = load_bf8(WN + (int64_t)(16 * c + lr) * GF + k0 + lk);
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int c = 0; c < GK / 16; c++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            dX[(row0 + 4 * (lane >> 4) + i) * GK + 16 * c + lr] = acc[c][i];
}

// weight-gradient partials X^T dZ over GDW_CHUNK points per workgroup: [320 x 64].
// Each 32-point step is staged transposed in LDS; wave w owns input rows 80w .. 80w+79.
#define XT_LD 40  // 32 points + 8 pad (bf16), 80-byte rows keep 16-byte alignment
__global__ __launch_bounds__(256) void k_gnn_gemm_dw(const uint16_t *X, const uint16_t *dZ,
                                                     float *part)
{
    __shared__ __attribute__((aligned(16))) uint16_t XT[GK * XT_LD];
    __shared__ __attribute__((aligned(16))) uint16_t ZT[GF * XT_LD];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = lane & 15, lk = (lane >> 4) * 8;
    f32x4 acc[5][4];
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int64_t pbase = (int64_t)blockIdx.x * GDW_CHUNK;
    for (int s = 0; s < GDW_CHUNK; s += 32) {
        const int64_t p0 = pbase + s;
        // stage: 32 x 320 of X and 32 x 64 of dZ, transposed
        for (int e = threadIdx.x; e < 32 * GK; e += 256) {
            const int pp = e / GK, k = e - pp * GK;
            XT[k * XT_LD + pp] = X[(p0 + pp) * GK + k];
        }
        for (int e = threadIdx.x; e < 32 * GF; e += 256) {
            const int pp = e / GF, k = e - pp * GF;
            ZT[k * XT_LD + pp] = dZ[(p0 + pp) * GF + k];
        }
        __syncthreads();
        bf16x8 b[4];
#pragma unroll
        for (int c = 0; c < 4; c++) b[c] = load_bf8(ZT + (16 * c + lr) * XT_LD + lk);
#pragma unroll
        for (int r = 0; r < 5; r++) {
            const bf16x8 a = load_bf8(XT + (80 * w + 16 * r + lr) * XT_LD + lk);
#pragma unroll
            for (int c = 0; c < 4; c++)
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[c], acc[r][c], 0, 0, 0);
        }
        __syncthreads();
    }
    float *out = part + (int64_t)blockIdx.x * (GK * GF);
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                out[(80 * w + 16 * r + 4 * (lane >> 4) + i) * GF + 16 * c + lr] = acc[r][c][i];
}

// gradient of the layer input H: direct, residual, aggregation adjoint and pooling adjoints
__global__ __launch_bounds__(256) void k_gnn_hbwd(int64_t n, int64_t npad, const float *dH,
                                                  const float *dX, const int64_t *near_ptr,
                                                  const int32_t *near, const float *S1,
                                                  const float *S2, const float *Sg, float *dHin)
{
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int f = threadIdx.x & 63;
    if (p >= npad) return;
    if (p >= n) {
        dHin[p * GF + f] = 0.f;
        return;
    }
    float g = dH[p * GF + f] + dX[p * GK + f];
    const int64_t j0 = near_ptr[p], j1 = near_ptr[p + 1];
    int64_t jj = j0;
    for (; jj + 4 <= j1; jj += 4) {
        int64_t q[4];
        float v[4], d[4];
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = near[jj + u];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            v[u] = dX[q[u] * GK + GF + f];
            d[u] = (float)(near_ptr[q[u] + 1] - near_ptr[q[u]]);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) g += v[u] / d[u];
    }
    for (; jj < j1; jj++) {
        const int64_t j = near[jj];
        const float dj = (float)(near_ptr[j + 1] - near_ptr[j]);
        g += dX[j * GK + GF + f] / dj;
    }
    const int64_t b1 = p / GB1, b2 = p / GB2;
    // S1/S2/Sg hold sums of dX columns 128.. (group 0 = P, 1 = Q, 2 = G)
    g += S1[b1 * (GF * 3) + f] / blk_count(n, b1, GB1);
    g += S2[b2 * (GF * 3) + GF + f] / blk_count(n, b2, GB2);
    g += Sg[2 * GF + f] / (float)n;
    dHin[p * GF + f] = g;
}

// ----------------------------------------------------------------------------------------
// optimiser and weight copies
// ----------------------------------------------------------------------------------------
__global__ void k_gnn_adam(int64_t np, float *theta, const float *g, float *m, float *v, float lr,
                           float b1, float b2, float c1, float c2, float eps)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    theta[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// bf16 copies of W_l: out-major WT[o][i] and in-major WN[i][o]
__global__ void k_gnn_wcopy(const float *theta, uint16_t *wT0, uint16_t *wN0, uint16_t *wT1,
                            uint16_t *wN1, uint16_t *wT2, uint16_t *wN2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= GL * GK * GF) return;
    const int l = e / (GK * GF), r = e - l * (GK * GF);
    const int i = r / GF, o = r - i * GF;
    const uint16_t h = f2bf(theta[OFF_W(l) + r]);
    uint16_t *wT = l == 0 ? wT0 : (l == 1 ? wT1 : wT2);
    uint16_t *wN = l == 0 ? wN0 : (l == 1 ? wN1 : wN2);
    wT[o * GK + i] = h;
    wN[i * GF + o] = h;
}

// training sample: y = a Y[i] + b Y[j], x likewise (float pairs -> FP64)
__global__ void k_gnn_mix(int64_t m, const float *Yi, const float *Xi, const float *Yj,
                          const float *Xj, double a, double b, double *y, double *x)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    double yy = a * (double)Yi[k], xx = a * (double)Xi[k];
    if (Yj) {
        yy += b * (double)Yj[k];
        xx += b * (double)Xj[k];
    }
    y[k] = yy;
    x[k] = xx;
}

__global__ void k_gnn_store_pair(int64_t m, const double *y, const double *x, float *Y, float *X)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    Y[k] = (float)y[k];
    X[k] = (float)x[k];
}

// sum over free dofs of (x - q)^2 (output-scale estimate), block partials
__global__ __launch_bounds__(256) void k_gnn_xq(int64_t m, const double *x, const double *q,
                                                const uint8_t *cmask, double *part)
{
    __shared__ double sh[256];
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (k < m && !cmask[k]) {
        const double d = x[k] - q[k];
        s = d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ----------------------------------------------------------------------------------------
// host
// ----------------------------------------------------------------------------------------
void pdx_gnn_free(pdx_ctx *c)
{
    pdx_gnn *g = c->gnn;
    if (!g) return;
    std::vector<void *> ps = {g->near_ptr, g->near, g->stat, g->theta, g->grad, g->am, g->av,
                              g->F, g->Z0, g->O, g->q, g->S1, g->S2, g->Sg, g->scal, g->dpart,
                              g->dH, g->dH2, g->dZf, g->dO, g->dZ, g->dX, g->wpart, g->Ky,
                              g->Kx, g->Ny, g->Nx, g->ys, g->xs};
    for (int l = 0; l < GL; l++) {
        ps.push_back(g->wT[l]);
        ps.push_back(g->wN[l]);
        ps.push_back(g->X[l]);
        ps.push_back(g->Z[l]);
    }
    for (int l = 0; l <= GL; l++) ps.push_back(g->H[l]);
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete g;
    c->gnn = nullptr;
}

static pdx_status gnn_alloc_params(pdx_ctx *c, pdx_gnn *g)
{
    if (g->theta) return PDX_OK;
    PDX_TRY(pdx_alloc(c, &g->theta, G_NPARAM));
    PDX_TRY(pdx_alloc(c, &g->grad, G_NPARAM));
    PDX_TRY(pdx_alloc(c, &g->am, G_NPARAM));
    PDX_TRY(pdx_alloc(c, &g->av, G_NPARAM));
    for (int l = 0; l < GL; l++) {
        PDX_TRY(pdx_alloc(c, &g->wT[l], GK * GF));
        PDX_TRY(pdx_alloc(c, &g->wN[l], GK * GF));
    }
    PDX_TRY(pdx_alloc(c, &g->scal, SC_N));
    PDX_HIP(c, hipMemset(g->scal, 0, sizeof(double) * SC_N));
    return PDX_OK;
}

static pdx_status gnn_wcopy(pdx_ctx *c, pdx_gnn *g)
{
    hipLaunchKernelGGL(k_gnn_wcopy, g1(GL * GK * GF), dim3(256), 0, c->stream, g->theta, g->wT[0],
                       g->wN[0], g->wT[1], g->wN[1], g->wT[2], g->wN[2]);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

// He-style initial weights, zero decoder (the untrained network is block Jacobi)
static pdx_status gnn_init_params(pdx_ctx *c, pdx_gnn *g, uint64_t seed)
{
    std::vector<float> th(G_NPARAM, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    for (int k = 0; k < GFIN * GF; k++) th[OFF_WE + k] = nd(rng) * std::sqrt(2.f / GFIN);
    for (int l = 0; l < GL; l++)
        for (int k = 0; k < GK * GF; k++) th[OFF_W(l) + k] = 0.5f * nd(rng) * std::sqrt(2.f / GK);
    PDX_HIP(c, hipMemcpy(g->theta, th.data(), sizeof(float) * G_NPARAM, hipMemcpyHostToDevice));
    PDX_HIP(c, hipMemset(g->am, 0, sizeof(float) * G_NPARAM));
    PDX_HIP(c, hipMemset(g->av, 0, sizeof(float) * G_NPARAM));
    g->adam_t = 0;
    return gnn_wcopy(c, g);
}

// graph, static features and activation buffers (once per context)
static pdx_status gnn_setup(pdx_ctx *c)
{
    if (!c->gnn) c->gnn = new pdx_gnn();
    pdx_gnn *g = c->gnn;
    if (g->ready) return PDX_OK;
    if (c->kind != PDX_BB)
        return pdx_fail(c, PDX_ERR_UNSUPPORTED, "the learned preconditioner is built for BBMaterial");
    const int64_t n = c->n;
    g->n = n;
    g->npad = (n + GPAD - 1) / GPAD * GPAD;
    g->nb1 = (n + GB1 - 1) / GB1;
    g->nb2 = (n + GB2 - 1) / GB2;
    const double lmax = G_NEAR_FRAC * c->P.delta;
    int32_t *dcnt = nullptr;
    PDX_TRY(pdx_alloc(c, &dcnt, n));
    hipLaunchKernelGGL(k_near_count, g1(n), dim3(256), 0, c->stream, n, c->slice_ptr, c->cnt, c->L,
                       lmax, dcnt);
    PDX_HIP(c, hipGetLastError());
    std::vector<int32_t> nc(n), bc(n);
    PDX_HIP(c, hipMemcpyAsync(nc.data(), dcnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
    PDX_HIP(c, hipMemcpyAsync(bc.data(), c->cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    (void)hipFree(dcnt);
    std::vector<int64_t> ptr(n + 1, 0);
    for (int64_t p = 0; p < n; p++) ptr[p + 1] = ptr[p] + nc[p];
    g->n_near = ptr[n];
    PDX_TRY(pdx_alloc(c, &g->near_ptr, n + 1));
    PDX_TRY(pdx_alloc(c, &g->near, std::max<int64_t>(g->n_near, 1)));
    PDX_HIP(c, hipMemcpy(g->near_ptr, ptr.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_near_fill, g1(n), dim3(256), 0, c->stream, n, c->slice_ptr, c->cnt, c->bw,
                       c->L, lmax, g->near_ptr, g->near);
    PDX_HIP(c, hipGetLastError());
    // static features from the reference positions
    std::vector<double> Xr(4 * n);
    PDX_HIP(c, hipMemcpy(Xr.data(), c->Xr, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    int32_t nmax = 1, bmax = 1;
    for (int64_t p = 0; p < n; p++) {
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], Xr[4 * p + d]);
            hi[d] = std::max(hi[d], Xr[4 * p + d]);
        }
        nmax = std::max(nmax, nc[p]);
        bmax = std::max(bmax, bc[p]);
    }
    std::vector<float> st(5 * n);
    for (int64_t p = 0; p < n; p++) {
        for (int d = 0; d < 3; d++) {
            const double half = std::max(0.5 * (hi[d] - lo[d]), 1e-300);
            st[5 * p + d] = (float)((Xr[4 * p + d] - 0.5 * (hi[d] + lo[d])) / half);
        }
        st[5 * p + 3] = (float)nc[p] / (float)nmax;
        st[5 * p + 4] = (float)bc[p] / (float)bmax;
    }
    PDX_TRY(pdx_alloc(c, &g->stat, 5 * n));
    PDX_HIP(c, hipMemcpy(g->stat, st.data(), sizeof(float) * 5 * n, hipMemcpyHostToDevice));
    // activations
    const int64_t np = g->npad;
    PDX_TRY(pdx_alloc(c, &g->F, np * GFIN));
    PDX_TRY(pdx_alloc(c, &g->Z0, np * GF));
    for (int l = 0; l <= GL; l++) PDX_TRY(pdx_alloc(c, &g->H[l], np * GF));
    for (int l = 0; l < GL; l++) {
        PDX_TRY(pdx_alloc(c, &g->X[l], np * GK));
        PDX_TRY(pdx_alloc(c, &g->Z[l], np * GF));
    }
    PDX_TRY(pdx_alloc(c, &g->O, 3 * n));
    PDX_TRY(pdx_alloc(c, &g->q, 3 * n));
    PDX_TRY(pdx_alloc(c, &g->S1, g->nb1 * GF * 3));
    PDX_TRY(pdx_alloc(c, &g->S2, g->nb2 * GF * 3));
    PDX_TRY(pdx_alloc(c, &g->Sg, GF * 3));
    // partial sums: 3 per 256 points (q), 2 per 16 points (decoder), 1 per 256 dofs (x - q)
    const int64_t nblk = (3 * n + 255) / 256 + 1;
    PDX_TRY(pdx_alloc(c, &g->dpart, std::max<int64_t>(3 * nblk, 2 * ((n + 15) / 16) + 2)));
    PDX_TRY(pdx_alloc(c, &g->dH, np * GF));
    PDX_TRY(pdx_alloc(c, &g->dH2, np * GF));
    PDX_TRY(pdx_alloc(c, &g->dZf, np * GF));
    PDX_TRY(pdx_alloc(c, &g->dO, 3 * np));
    PDX_TRY(pdx_alloc(c, &g->dZ, np * GF));
    PDX_TRY(pdx_alloc(c, &g->dX, np * GK));
    g->wpart_len = std::max<int64_t>((np / GDW_CHUNK) * GK * GF, (np / 256 + 1) * (GK + 1) * GF);
    PDX_TRY(pdx_alloc(c, &g->wpart, g->wpart_len));
    PDX_TRY(pdx_alloc(c, &g->ys, 3 * n));
    PDX_TRY(pdx_alloc(c, &g->xs, 3 * n));
    PDX_TRY(gnn_alloc_params(c, g));
    g->ready = true;
    return PDX_OK;
}

// forward pass for the FP64 vector y (internal order).  TRAIN keeps the activations and
// accumulates the loss against x; otherwise z = M(y).
static pdx_status gnn_forward(pdx_ctx *c, const double *y, const double *x, double *z, bool train,
                              double inv_batch)
{
    pdx_gnn *g = c->gnn;
    pdx_nk *k = c->nk;
    const int64_t n = g->n, np = g->npad;
    const int nb = (int)((n + 255) / 256);
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_gnn_q, dim3(nb), dim3(256), 0, s, n, k->dinv, y, k->cmask, g->q, g->dpart);
    hipLaunchKernelGGL(k_gnn_q_final, dim3(1), dim3(256), 0, s, nb, g->dpart, g->scal);
    hipLaunchKernelGGL(k_gnn_encode, dim3((unsigned)(np / 4)), dim3(256), 0, s, n, np, g->q, y, k->cmask, g->stat,
                       g->scal, g->theta, train ? g->F : nullptr, train ? g->Z0 : nullptr, g->H[0]);
    PDX_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_gnn_pool1<uint16_t>, dim3((unsigned)g->nb1), dim3(4 * GF), 0, s, n, 1, g->H[0],
                       GF, 0, g->S1);
    for (int l = 0; l < GL; l++) {
        hipLaunchKernelGGL(k_gnn_pool2, dim3((unsigned)g->nb2), dim3(GF), 0, s, g->nb1, 1, g->S1, g->S2);
        hipLaunchKernelGGL(k_gnn_poolg, dim3(1), dim3(GF), 0, s, g->nb2, g->S2, g->Sg);
        hipLaunchKernelGGL(k_gnn_concat, dim3((unsigned)(np / 4)), dim3(256), 0, s, n, np, g->H[l],
                           g->near_ptr, g->near, g->S1, g->S2, g->Sg, g->X[l]);
        float *S1n = l + 1 < GL ? g->S1 : nullptr;  // pooled sums of H[l + 1] for the next layer
        if (train)
            hipLaunchKernelGGL(k_gnn_gemm_fwd<true>, dim3((unsigned)(np / 64)), dim3(256), 0, s, n,
                               g->X[l], g->wT[l], g->theta + OFF_B(l), g->H[l], g->H[l + 1], g->Z[l], S1n);
        else
            hipLaunchKernelGGL(k_gnn_gemm_fwd<false>, dim3((unsigned)(np / 64)), dim3(256), 0, s, n,
                               g->X[l], g->wT[l], g->theta + OFF_B(l), g->H[l], g->H[l + 1], nullptr, S1n);
        PDX_HIP(c, hipGetLastError());
    }
    if (train) {
        const int nbd = (int)((n + 15) / 16);
        hipLaunchKernelGGL(k_gnn_decode<true>, dim3(nbd), dim3(256), 0, s, n, g->H[GL], g->theta, g->q,
                           y, x, k->cmask, g->scal, nullptr, g->O, g->dpart);
        hipLaunchKernelGGL(k_gnn_loss_final, dim3(1), dim3(256), 0, s, nbd, g->dpart, g->scal, inv_batch);
    } else {
        hipLaunchKernelGGL(k_gnn_decode<false>, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, s, n, g->H[GL], g->theta, g->q,
                           y, nullptr, k->cmask, g->scal, z, nullptr, nullptr);
    }
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

// adds the gradient of the last training forward pass to g->grad
static pdx_status gnn_backward(pdx_ctx *c, const double *x)
{
    pdx_gnn *g = c->gnn;
    pdx_nk *k = c->nk;
    const int64_t n = g->n, np = g->npad;
    hipStream_t s = c->stream;
    const int64_t nch = np / 256;  // 256-point chunks
    hipLaunchKernelGGL(k_gnn_decode_bwd, dim3((unsigned)(np / 4)), dim3(256), 0, s, n, np, g->theta, g->O, g->q, x,
                       k->cmask, g->scal, g->dO, g->dH);
    // W_d, b_d
    hipLaunchKernelGGL(k_gnn_outer<uint16_t>, dim3((unsigned)nch), dim3(256), 0, s, np, (int64_t)256,
                       g->H[GL], GF, g->dO, 3, g->wpart);
    hipLaunchKernelGGL(k_gnn_reduce_add, g1((GF + 1) * 3), dim3(256), 0, s, nch, (int64_t)(GF + 1) * 3,
                       g->wpart, g->grad + OFF_WD);
    PDX_HIP(c, hipGetLastError());
    float *dHo = g->dH, *dHi = g->dH2;
    for (int l = GL - 1; l >= 0; l--) {
        hipLaunchKernelGGL(k_gnn_relu_bwd, dim3((unsigned)nch), dim3(256), 0, s, n, dHo, g->Z[l], g->dZ,
                           nullptr, g->wpart);
        hipLaunchKernelGGL(k_gnn_reduce_add, dim3(1), dim3(GF), 0, s, nch, (int64_t)GF, g->wpart,
                           g->grad + OFF_B(l));
        hipLaunchKernelGGL(k_gnn_gemm_dw, dim3((unsigned)(np / GDW_CHUNK)), dim3(256), 0, s, g->X[l],
                           g->dZ, g->wpart);
        hipLaunchKernelGGL(k_gnn_reduce_add, g1(GK * GF), dim3(256), 0, s, np / GDW_CHUNK,
                           (int64_t)GK * GF, g->wpart, g->grad + OFF_W(l));
        hipLaunchKernelGGL(k_gnn_gemm_dx, dim3((unsigned)(np / 64)), dim3(256), 0, s, g->dZ, g->wN[l],
                           g->dX);
        // pooled adjoints of dX columns 128..319
        hipLaunchKernelGGL(k_gnn_pool1<float>, dim3((unsigned)g->nb1), dim3(4 * 3 * GF), 0, s, n, 3, g->dX,
                           GK, 2 * GF, g->S1);
        hipLaunchKernelGGL(k_gnn_pool2, dim3((unsigned)g->nb2), dim3(3 * GF), 0, s, g->nb1, 3, g->S1, g->S2);
        hipLaunchKernelGGL(k_gnn_poolg, dim3(1), dim3(3 * GF), 0, s, g->nb2, g->S2, g->Sg);
        hipLaunchKernelGGL(k_gnn_hbwd, dim3((unsigned)(np / 4)), dim3(256), 0, s, n, np, dHo, g->dX,
                           g->near_ptr, g->near, g->S1, g->S2, g->Sg, dHi);
        PDX_HIP(c, hipGetLastError());
        std::swap(dHo, dHi);
    }
    // encoder
    hipLaunchKernelGGL(k_gnn_relu_bwd, dim3((unsigned)nch), dim3(256), 0, s, n, dHo, g->Z0, nullptr,
                       g->dZf, nullptr);
    hipLaunchKernelGGL(k_gnn_outer<float>, dim3((unsigned)nch), dim3(256), 0, s, np, (int64_t)256, g->F,
                       GFIN, g->dZf, GF, g->wpart);
    hipLaunchKernelGGL(k_gnn_reduce_add, g1((GFIN + 1) * GF), dim3(256), 0, s, nch,
                       (int64_t)(GFIN + 1) * GF, g->wpart, g->grad + OFF_WE);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

static pdx_status gnn_set_c(pdx_ctx *c, double cval)
{
    pdx_gnn *g = c->gnn;
    g->c_out = cval;
    PDX_HIP(c, hipMemcpyAsync(g->scal + SC_C, &g->c_out, sizeof(double), hipMemcpyHostToDevice, c->stream));
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    return PDX_OK;
}

// ---------------------------------------------------------------------------- NK hooks
pdx_status pdx_gnn_prepare(pdx_ctx *c)
{
    PDX_TRY(gnn_setup(c));
    if (!c->gnn->trained)
        return pdx_fail(c, PDX_ERR_STATE, "the learned preconditioner is not trained (pdx_gnn_train)");
    return PDX_OK;
}

// the learned preconditioner of the Krylov solver: two-level coarse correction plus the
// network's block-Jacobi-based output, z = P A_c^-1 P^T v + D^-1 v + c s O(v)
pdx_status pdx_nk_two_level(pdx_ctx *c, const double *v, double *z);

pdx_status pdx_gnn_apply(pdx_ctx *c, const double *v, double *z)
{
    // a network that failed validation is not applied: the two-level preconditioner alone
    if (!c->gnn->use_net && c->nk->coarse_ready) return pdx_nk_two_level(c, v, z);
    PDX_TRY(gnn_forward(c, v, nullptr, z, false, 0.0));
    if (c->nk->coarse_ready) PDX_TRY(pdx_nk_coarse_add(c, v, z, 1.0));
    return PDX_OK;
}

static pdx_status gnn_ensure_pool(pdx_ctx *c)
{
    pdx_gnn *g = c->gnn;
    const int64_t m = 3 * c->n;
    if (!g->Ny) {
        g->cap_n = 256;
        PDX_TRY(pdx_alloc(c, &g->Ny, g->cap_n * m));
        PDX_TRY(pdx_alloc(c, &g->Nx, g->cap_n * m));
    }
    if (!g->Ky && g->cap_k > 0) {
        PDX_TRY(pdx_alloc(c, &g->Ky, g->cap_k * m));
        PDX_TRY(pdx_alloc(c, &g->Kx, g->cap_k * m));
    }
    return PDX_OK;
}

// Newton pair (y = -r, x = du)
pdx_status pdx_gnn_record(pdx_ctx *c, const double *rhs, const double *du)
{
    pdx_gnn *g = c->gnn;
    if (!g || !g->collecting) return PDX_OK;
    PDX_TRY(gnn_ensure_pool(c));
    const int64_t m = 3 * c->n;
    const int64_t slot = g->n_n < g->cap_n ? g->n_n : (int64_t)(g->rng() % (uint64_t)g->cap_n);
    hipLaunchKernelGGL(k_gnn_store_pair, g1(m), dim3(256), 0, c->stream, m, rhs, du, g->Ny + slot * m,
                       g->Nx + slot * m);
    PDX_HIP(c, hipGetLastError());
    if (g->n_n < g->cap_n) g->n_n++;
    return PDX_OK;
}

// Krylov pair (y = J z, x = z), reservoir-sampled into cap_k slots
pdx_status pdx_gnn_record_krylov(pdx_ctx *c, const double *z, const double *jz)
{
    pdx_gnn *g = c->gnn;
    if (!g || !g->collecting || g->cap_k <= 0) return PDX_OK;
    PDX_TRY(gnn_ensure_pool(c));
    const int64_t m = 3 * c->n;
    int64_t slot = -1;
    if (g->seen_k < g->cap_k) slot = g->seen_k;
    else {
        const uint64_t r = g->rng() % (uint64_t)(g->seen_k + 1);
        if ((int64_t)r < g->cap_k) slot = (int64_t)r;
    }
    g->seen_k++;
    if (slot < 0) return PDX_OK;
    hipLaunchKernelGGL(k_gnn_store_pair, g1(m), dim3(256), 0, c->stream, m, jz, z, g->Ky + slot * m,
                       g->Kx + slot * m);
    PDX_HIP(c, hipGetLastError());
    g->n_k = std::min(g->seen_k, g->cap_k);
    return PDX_OK;
}

bool pdx_gnn_collecting(pdx_ctx *c) { return c->gnn && c->gnn->collecting; }
void pdx_gnn_disable(pdx_ctx *c)
{
    if (c->gnn) c->gnn->use_net = false;
}
bool pdx_gnn_active(pdx_ctx *c) { return c->gnn && c->gnn->trained && c->gnn->use_net; }

// ---------------------------------------------------------------------------- C ABI
static pdx_status gnn_ready_for_io(pdx_ctx *c)
{
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (!c->nk || !c->nk->cmask) return pdx_fail(c, PDX_ERR_STATE, "pdx_set_displacement_bcs first");
    PDX_HIP(c, hipSetDevice(c->device));
    return gnn_setup(c);
}

extern "C" pdx_status pdx_gnn_collect(pdx_ctx *c, int32_t enable, int64_t max_krylov_pairs)
{
    if (!c) return PDX_ERR_ARG;
    if (max_krylov_pairs < 0) return pdx_fail(c, PDX_ERR_ARG, "max_krylov_pairs < 0");
    PDX_TRY(gnn_ready_for_io(c));
    pdx_gnn *g = c->gnn;
    if (enable && !g->collecting && g->Ky == nullptr) g->cap_k = max_krylov_pairs;
    g->collecting = enable != 0;
    return PDX_OK;
}

extern "C" pdx_status pdx_gnn_pairs(pdx_ctx *c, int64_t *n_newton, int64_t *n_krylov)
{
    if (!c) return PDX_ERR_ARG;
    pdx_gnn *g = c->gnn;
    if (n_newton) *n_newton = g ? g->n_n : 0;
    if (n_krylov) *n_krylov = g ? g->n_k : 0;
    return PDX_OK;
}

static pdx_status gnn_build_dinv(pdx_ctx *c);

// one sample into ys/xs: Newton pairs, Krylov pairs, and combinations of two
static pdx_status gnn_draw(pdx_ctx *c, std::mt19937_64 &rng)
{
    pdx_gnn *g = c->gnn;
    const int64_t m = 3 * c->n;
    std::uniform_real_distribution<double> ud(0.0, 1.0);
    std::normal_distribution<double> nd(0.0, 1.0);
    auto pick = [&](const float *&Y, const float *&X) {
        const bool newton = g->n_k == 0 || (g->n_n > 0 && ud(rng) < 0.5);
        if (newton) {
            const int64_t i = (int64_t)(rng() % (uint64_t)g->n_n);
            Y = g->Ny + i * m;
            X = g->Nx + i * m;
        } else {
            const int64_t i = (int64_t)(rng() % (uint64_t)(g->n_k - g->n_k_val));
            Y = g->Ky + i * m;
            X = g->Kx + i * m;
        }
    };
    const float *Y1, *X1, *Y2 = nullptr, *X2 = nullptr;
    pick(Y1, X1);
    double a = 1.0, b = 0.0;
    if (ud(rng) < 0.5) {
        pick(Y2, X2);
        a = nd(rng);
        b = nd(rng);
    }
    hipLaunchKernelGGL(k_gnn_mix, g1(m), dim3(256), 0, c->stream, m, Y1, X1, Y2, X2, a, b, g->ys, g->xs);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

// output scale c = median over samples of rms_free(x' - D^-1 y) / rms_free(D^-1 y), with
// x' = x - P A_c^-1 P^T y when the coarse space is built
static pdx_status gnn_estimate_c(pdx_ctx *c, std::mt19937_64 &rng)
{
    pdx_gnn *g = c->gnn;
    const int64_t m = 3 * c->n, n = c->n;
    const int nb = (int)((n + 255) / 256), nbm = (int)((m + 255) / 256);
    std::vector<double> ratios;
    const int ns = (int)std::min<int64_t>(32, g->n_n + g->n_k);
    for (int k = 0; k < ns; k++) {
        PDX_TRY(gnn_draw(c, rng));
        if (c->nk->coarse_ready) PDX_TRY(pdx_nk_coarse_add(c, g->ys, g->xs, -1.0));
        hipLaunchKernelGGL(k_gnn_q, dim3(nb), dim3(256), 0, c->stream, n, c->nk->dinv, g->ys,
                           c->nk->cmask, g->q, g->dpart);
        hipLaunchKernelGGL(k_gnn_q_final, dim3(1), dim3(256), 0, c->stream, nb, g->dpart, g->scal);
        double sc[3];
        PDX_HIP(c, hipMemcpyAsync(sc, g->scal, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
        hipLaunchKernelGGL(k_gnn_xq, dim3(nbm), dim3(256), 0, c->stream, m, g->xs, g->q, c->nk->cmask,
                           g->dpart);
        std::vector<double> part(nbm);
        PDX_HIP(c, hipMemcpyAsync(part.data(), g->dpart, sizeof(double) * nbm, hipMemcpyDeviceToHost,
                                  c->stream));
        PDX_HIP(c, hipStreamSynchronize(c->stream));
        double xq = 0.0;
        for (double v : part) xq += v;
        const double nf = std::max(sc[SC_NFREE], 1.0);
        const double r = std::sqrt(xq / nf) / sc[SC_S];
        if (std::isfinite(r) && r > 0) ratios.push_back(r);
    }
    double cval = 1.0;
    if (!ratios.empty()) {
        std::sort(ratios.begin(), ratios.end());
        cval = ratios[ratios.size() / 2];
    }
    return gnn_set_c(c, cval);
}

extern "C" pdx_status pdx_gnn_train(pdx_ctx *c, int32_t steps, int32_t batch, double lr,
                                    uint64_t seed, double *loss_out)
{
    if (!c) return PDX_ERR_ARG;
    if (steps < 0 || batch <= 0 || !(lr > 0)) return pdx_fail(c, PDX_ERR_ARG, "invalid training options");
    PDX_TRY(gnn_ready_for_io(c));
    pdx_gnn *g = c->gnn;
    if (g->n_n + g->n_k == 0) return pdx_fail(c, PDX_ERR_STATE, "no training pairs (pdx_gnn_collect)");
    PDX_TRY(gnn_build_dinv(c));
    // the network learns what the two-level preconditioner leaves: x - P A_c^-1 P^T y
    PDX_TRY(pdx_nk_build_coarse(c));
    std::mt19937_64 rng(seed);
    // hold out the last tenth of the Krylov pairs (if there are enough) for validation
    g->n_k_val = g->n_k >= 20 ? g->n_k / 10 : 0;
    if (g->n_n == 0 && g->n_k - g->n_k_val <= 0) g->n_k_val = 0;
    if (!g->trained) PDX_TRY(gnn_init_params(c, g, seed));
    PDX_TRY(gnn_estimate_c(c, rng));
    const float b1 = 0.9f, b2 = 0.999f;
    double last = 0.0;
    for (int it = 0; it < steps; it++) {
        PDX_HIP(c, hipMemsetAsync(g->grad, 0, sizeof(float) * G_NPARAM, c->stream));
        PDX_HIP(c, hipMemsetAsync(g->scal + SC_LOSS, 0, sizeof(double), c->stream));
        for (int b = 0; b < batch; b++) {
            PDX_TRY(gnn_draw(c, rng));
            PDX_TRY(pdx_nk_coarse_add(c, g->ys, g->xs, -1.0));
            PDX_TRY(gnn_forward(c, g->ys, g->xs, nullptr, true, 1.0 / batch));
            PDX_TRY(gnn_backward(c, g->xs));
        }
        g->adam_t++;
        const float c1 = 1.f - std::pow(b1, (float)g->adam_t), c2 = 1.f - std::pow(b2, (float)g->adam_t);
        // cosine decay of the step size
        const float lrk = (float)(lr * 0.5 * (1.0 + std::cos(M_PI * it / std::max(steps, 1))));
        hipLaunchKernelGGL(k_gnn_adam, g1(G_NPARAM), dim3(256), 0, c->stream, (int64_t)G_NPARAM, g->theta,
                           g->grad, g->am, g->av, lrk, b1, b2, c1, c2, 1e-8f);
        PDX_TRY(gnn_wcopy(c, g));
        if (it == steps - 1 || (it % 50) == 49) {
            PDX_HIP(c, hipMemcpyAsync(&last, g->scal + SC_LOSS, sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream));
            PDX_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    g->trained = true;
    if (loss_out) *loss_out = last;
    // held-out loss ||c s O - e||^2 / ||e||^2 of the two-level error e (1 = no gain)
    g->val_loss = last;
    if (g->n_k_val > 0) {
        const int64_t m = 3 * c->n;
        PDX_HIP(c, hipMemsetAsync(g->scal + SC_LOSS, 0, sizeof(double), c->stream));
        for (int64_t i = g->n_k - g->n_k_val; i < g->n_k; i++) {
            hipLaunchKernelGGL(k_gnn_mix, g1(m), dim3(256), 0, c->stream, m, g->Ky + i * m, g->Kx + i * m,
                               (const float *)nullptr, (const float *)nullptr, 1.0, 0.0, g->ys, g->xs);
            PDX_TRY(pdx_nk_coarse_add(c, g->ys, g->xs, -1.0));
            PDX_TRY(gnn_forward(c, g->ys, g->xs, nullptr, true, 1.0 / (double)g->n_k_val));
        }
        PDX_HIP(c, hipMemcpyAsync(&g->val_loss, g->scal + SC_LOSS, sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
        PDX_HIP(c, hipStreamSynchronize(c->stream));
    }
    g->use_net = g->val_loss < PDX_GNN_GATE;
    return PDX_OK;
}

extern "C" pdx_status pdx_gnn_gate(pdx_ctx *c, double threshold, double *val_loss, int32_t *use_network)
{
    if (!c) return PDX_ERR_ARG;
    PDX_TRY(gnn_ready_for_io(c));
    pdx_gnn *g = c->gnn;
    if (!g->trained) return pdx_fail(c, PDX_ERR_STATE, "the network is not trained (pdx_gnn_train)");
    if (threshold == threshold) g->use_net = g->val_loss < threshold;  // NaN: only report
    if (val_loss) *val_loss = g->val_loss;
    if (use_network) *use_network = g->use_net ? 1 : 0;
    return PDX_OK;
}

// point-block Jacobi at the current state (pdx_nk.hip)
pdx_status pdx_nk_build_jacobi(pdx_ctx *c);
static pdx_status gnn_build_dinv(pdx_ctx *c) { return pdx_nk_build_jacobi(c); }

extern "C" pdx_status pdx_gnn_params(pdx_ctx *c, float *theta_out, const float *theta_in,
                                     double *out_scale)
{
    if (!c) return PDX_ERR_ARG;
    PDX_TRY(gnn_ready_for_io(c));
    pdx_gnn *g = c->gnn;
    if (theta_in) {
        PDX_HIP(c, hipMemcpy(g->theta, theta_in, sizeof(float) * G_NPARAM, hipMemcpyHostToDevice));
        PDX_HIP(c, hipMemset(g->am, 0, sizeof(float) * G_NPARAM));
        PDX_HIP(c, hipMemset(g->av, 0, sizeof(float) * G_NPARAM));
        g->adam_t = 0;
        PDX_TRY(gnn_wcopy(c, g));
        if (out_scale) PDX_TRY(gnn_set_c(c, *out_scale));
        g->trained = true;
    } else if (out_scale) {
        *out_scale = g->c_out;
    }
    if (theta_out) PDX_HIP(c, hipMemcpy(theta_out, g->theta, sizeof(float) * G_NPARAM, hipMemcpyDeviceToHost));
    PDX_HIP(c, hipDeviceSynchronize());
    return PDX_OK;
}

extern "C" int64_t pdx_gnn_n_params(void) { return G_NPARAM; }

extern "C" pdx_status pdx_gnn_graph(pdx_ctx *c, int64_t *n_near, int64_t *perm, int64_t *near_ptr,
                                    int32_t *near, float *stat, double *dinv, uint8_t *cmask)
{
    if (!c) return PDX_ERR_ARG;
    PDX_TRY(gnn_ready_for_io(c));
    pdx_gnn *g = c->gnn;
    const int64_t n = c->n;
    if (n_near) *n_near = g->n_near;
    if (perm) {
        std::vector<int64_t> pm(c->perm.begin(), c->perm.end());
        std::memcpy(perm, pm.data(), sizeof(int64_t) * n);
    }
    if (near_ptr) PDX_HIP(c, hipMemcpy(near_ptr, g->near_ptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    if (near) PDX_HIP(c, hipMemcpy(near, g->near, sizeof(int32_t) * g->n_near, hipMemcpyDeviceToHost));
    if (stat) PDX_HIP(c, hipMemcpy(stat, g->stat, sizeof(float) * 5 * n, hipMemcpyDeviceToHost));
    if (dinv) {
        PDX_TRY(gnn_build_dinv(c));
        PDX_HIP(c, hipMemcpy(dinv, c->nk->dinv, sizeof(double) * 9 * n, hipMemcpyDeviceToHost));
    }
    if (cmask) PDX_HIP(c, hipMemcpy(cmask, c->nk->cmask, 3 * n, hipMemcpyDeviceToHost));
    return PDX_OK;
}

extern "C" pdx_status pdx_gnn_apply_vec(pdx_ctx *c, const double *y, double *z)
{
    if (!c || !y || !z) return PDX_ERR_ARG;
    PDX_TRY(gnn_ready_for_io(c));
    const int64_t m = 3 * c->n;
    pdx_gnn *g = c->gnn;
    PDX_TRY(gnn_build_dinv(c));
    PDX_HIP(c, hipMemcpy(g->ys, y, sizeof(double) * m, hipMemcpyHostToDevice));
    PDX_TRY(gnn_forward(c, g->ys, nullptr, g->xs, false, 0.0));
    PDX_HIP(c, hipMemcpy(z, g->xs, sizeof(double) * m, hipMemcpyDeviceToHost));
    return PDX_OK;
}

extern "C" pdx_status pdx_gnn_loss_grad(pdx_ctx *c, int32_t n_samples, const double *Y,
                                        const double *X, double *loss, float *grad)
{
    if (!c || n_samples <= 0 || !Y || !X) return PDX_ERR_ARG;
    PDX_TRY(gnn_ready_for_io(c));
    pdx_gnn *g = c->gnn;
    const int64_t m = 3 * c->n;
    PDX_TRY(gnn_build_dinv(c));
    PDX_HIP(c, hipMemsetAsync(g->grad, 0, sizeof(float) * G_NPARAM, c->stream));
    PDX_HIP(c, hipMemsetAsync(g->scal + SC_LOSS, 0, sizeof(double), c->stream));
    for (int s = 0; s < n_samples; s++) {
        PDX_HIP(c, hipMemcpyAsync(g->ys, Y + s * m, sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
        PDX_HIP(c, hipMemcpyAsync(g->xs, X + s * m, sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
        PDX_TRY(gnn_forward(c, g->ys, g->xs, nullptr, true, 1.0 / n_samples));
        PDX_TRY(gnn_backward(c, g->xs));
        PDX_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (loss) PDX_HIP(c, hipMemcpy(loss, g->scal + SC_LOSS, sizeof(double), hipMemcpyDeviceToHost));
    if (grad) PDX_HIP(c, hipMemcpy(grad, g->grad, sizeof(float) * G_NPARAM, hipMemcpyDeviceToHost));
    return PDX_OK;
}
