// Learned preconditioner of the device Newton-Krylov solver: a message-passing network on
// the bond graph that generates the coarse space of the two-level preconditioner.  It is
// evaluated with BF16 MFMA (FP32 accumulation) and trained on the GPU by its own backward
// kernels and Adam.  No counterpart in the reference, which calls IterativeSolvers.gmres!
// without a preconditioner (src/time_solvers/newton_krylov.jl:448-450); the coarse space it
// builds plugs into the flexible GMRES of pdx_nk.hip as a right preconditioner, with the
// two-level preconditioner of the translations alone as fallback.
//
// Network, per point i (internal Morton order), theta in float32 (master copy):
//   F_i (16)  = [normalised X (3), X relative to its aggregate (3), constrained mask (3),
//               near-degree ratio, bond-count ratio, squares of the relative X (3), 1, 0]
//   H_i       = relu(W_e^T F_i + b_e)                                     16 -> 64
//   3 layers:   A_i = mean_{j in N(i)} H_j   (near bonds: L <= 0.6 delta, symmetric)
//               P_i, Q_i, G = means of H over the 64-point and 4096-point blocks of i and
//                             over the body
//               Z_i = W_l^T [H_i A_i P_i Q_i G] + b_l                     320 -> 64 (MFMA)
//               H_i = H_i + relu(Z_i)
//   O_i       = W_d^T H_i + b_d                                           64 -> 3 K
//   fields    phi_i[f] = (O_i[3f], O_i[3f+1], O_i[3f+2]),  f < K = 8
// The inputs depend only on the reference configuration, so the fields are computed once
// after training.  The coarse space of aggregate I is spanned by the three translations
// and the K fields restricted to I (free dofs); pdx_nk.hip assembles its Galerkin operator.
// Training samples are the errors e = z - M_T(J z) the two-level preconditioner of the
// translations leaves on the Krylov pairs (J z, z) of the solves it preconditions.  Loss:
//   (1/S) sum_s sum_I min_c ( |e_sI - B_I c|^2 + lambda_I |c|^2 ) / |e_s|^2,
// B_I = [translations | fields] on I's free dofs, lambda_I = 1e-6 tr(B_I^T B_I) / (3 + K);
// by the envelope theorem d/dB_I = -2 (e_sI - B_I c) c^T / (S |e_s|^2).
// H and the concatenated layer inputs are stored in bfloat16; every sum is float32.
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
#define GKF 8          // learned fields
#define GNO (3 * GKF)  // decoder outputs
#define GNB (3 + GKF)  // coarse fields per aggregate: translations + learned
#define GB1 64         // first pooling block
#define GB2 4096       // second pooling block
#define GPAD 512       // row padding of every activation buffer
#define GDW_CHUNK 512  // points per workgroup of the weight-gradient GEMM
#define G_NEAR_FRAC 0.6
#define G_AGG_MAX 512  // aggregate size the loss kernel stages in LDS
#define G_RIDGE 1e-6

// flat parameter layout (float32), identical to the CPU oracle's
#define OFF_WE 0
#define OFF_BE (OFF_WE + GFIN * GF)
#define OFF_W(l) (OFF_BE + GF + (l) * (GK * GF + GF))
#define OFF_B(l) (OFF_W(l) + GK * GF)
#define OFF_WD (OFF_BE + GF + GL * (GK * GF + GF))
#define OFF_BD (OFF_WD + GF * GNO)
#define G_NPARAM (OFF_BD + GNO)

// held-out loss ratio (fields / translations alone) below which the learned coarse space
// is used (1 = the fields explain nothing the translations do not)
#define PDX_GNN_GATE 0.95

struct pdx_gnn {
    int64_t n = 0, npad = 0, nb1 = 0, nb2 = 0;
    int agg = 0;
    int64_t nagg = 0;
    // graph and static features
    bool ready = false;
    int64_t n_near = 0;
    int64_t *near_ptr = nullptr;
    int32_t *near = nullptr;
    // parameters
    bool trained = false;
    bool have_fields = false;   // O holds the fields of the current parameters
    int64_t n_val = 0;          // held-out samples (the last tenth)
    double val_loss = -1.0;     // held-out loss ratio, fields / translations alone
    double val_loss_t = -1.0;   // held-out loss of the translations alone
    bool use_net = true;
    float *theta = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
    uint16_t *wT[GL] = {}, *wN[GL] = {};  // bf16: out-major [64][320], in-major [320][64]
    int64_t adam_t = 0;
    // activations (npad rows)
    float *F = nullptr, *Z0 = nullptr;
    uint16_t *H[GL + 1] = {};
    uint16_t *X[GL] = {};
    float *Z[GL] = {};
    float *O = nullptr;         // npad x GNO: the fields
    float *S1 = nullptr, *S2 = nullptr, *Sg = nullptr;  // pooled sums, 192 columns
    double *lpart = nullptr;    // loss partial per aggregate
    int32_t *sidx = nullptr;    // sample indices of a training step
    // backward
    float *dH = nullptr, *dH2 = nullptr, *dZf = nullptr, *dO = nullptr;
    uint16_t *dZ = nullptr;
    float *dX = nullptr;
    float *wpart = nullptr;     // weight-gradient partials
    int64_t wpart_len = 0;
    // error samples (float32, internal order) and their 1 / |e|^2 over free dofs
    bool collecting = false;
    int64_t cap = 0, n_s = 0, seen = 0;
    float *E = nullptr;
    float *einv = nullptr;
    double *tmp = nullptr, *tmp2 = nullptr;  // 3n
    std::mt19937_64 rng{12345};
};

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
// workgroups of k_gnn_reduce_add for len outputs (16 per workgroup)
static inline dim3 g16(int64_t len) { return dim3((unsigned)((len + 15) / 16)); }

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
// encoder: H0_i = relu(W_e^T F_i + b_e); one wavefront per point, lane = feature (the
// inputs are broadcast loads, the stores of H0 coalesced); Z0 kept for training
__global__ __launch_bounds__(256) void k_gnn_encode(int64_t n, int64_t npad, const float *Fi,
                                                    const float *theta, float *Z0, uint16_t *H0)
{
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int f = threadIdx.x & 63;
    if (p >= npad) return;
    if (p >= n) {
        H0[p * GF + f] = 0;
        if (Z0) Z0[p * GF + f] = 0.f;
        return;
    }
    float z = theta[OFF_BE + f];
#pragma unroll
    for (int k = 0; k < GFIN; k++) z += Fi[p * GFIN + k] * theta[OFF_WE + k * GF + f];
    if (Z0) Z0[p * GF + f] = z;
    H0[p * GF + f] = f2bf(fmaxf(z, 0.f));
}

// decoder O = W_d^T H + b_d (64 -> GNO): ten points per workgroup, a thread per output
__global__ __launch_bounds__(256) void k_gnn_decode(int64_t n, int64_t npad, const uint16_t *H,
                                                    const float *theta, float *O)
{
    const int pl = threadIdx.x / GNO, o = threadIdx.x % GNO;
    const int64_t p = (int64_t)blockIdx.x * (256 / GNO) + pl;
    if (pl >= 256 / GNO || p >= npad) return;
    float s = 0.f;
    if (p < n) {
        s = theta[OFF_BD + o];
        for (int f = 0; f < GF; f++) s += bf2f(H[p * GF + f]) * theta[OFF_WD + f * GNO + o];
    }
    O[p * GNO + o] = s;
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

// ----------------------------------------------------------------------------------------
// loss: per aggregate, the ridge least-squares fit of every error sample by the coarse
// fields, and the gradient with respect to the fields.  One workgroup per aggregate; the
// aggregate's basis B (rows = its 3 |I| dofs, columns = 3 translations + GKF fields, zero on
// constrained dofs) lives in LDS, the Gram matrix is factored once, and the samples are fitted
// in chunks of G_LSQ_CHUNK:
//   1. B^T e of every sample of the chunk: 16 threads per sample, each over every 16th row,
//      folded with lane shuffles;
//   2. the coefficients c = (B^T B + lambda I)^-1 B^T e, one thread per sample (Cholesky
//      substitution in FP64);
//   3. the residuals e - B c: each thread over its own rows and all samples of the chunk,
//      accumulating the gradient of its rows in registers.
// The ridge term's gradient, 2 ridge / (3 + K) |c|^2 B per sample, is applied once at the end
// with the chunk sums of w |c|^2.
// ----------------------------------------------------------------------------------------
#define G_LSQ_CHUNK 16
#define G_LSQ_ROWS ((3 * G_AGG_MAX + 255) / 256)  // rows per thread in the residual pass
__global__ __launch_bounds__(256) void k_gnn_lsq(int64_t n, int agg, const float *O,
                                                 const uint8_t *cmask, const float *E,
                                                 const float *einv, const int32_t *sidx,
                                                 int ns, int64_t m, int use_fields, float wscale,
                                                 float *dO, double *lpart)
{
    static_assert(G_LSQ_CHUNK * 16 == 256, "16 threads per sample of a chunk");
    __shared__ float B[3 * G_AGG_MAX][GNB];
    __shared__ double G[GNB][GNB], Lc[GNB][GNB];
    __shared__ float red[G_LSQ_CHUNK][GNB];
    __shared__ double cv[G_LSQ_CHUNK][GNB];
    __shared__ float wq[G_LSQ_CHUNK];
    __shared__ int64_t srow[G_LSQ_CHUNK];
    __shared__ double lsum[256];
    __shared__ double lam_sh, c2w_sh;
    const int64_t I = blockIdx.x;
    const int64_t p0 = I * agg, p1 = min(n, p0 + agg);
    const int nr = (int)(3 * (p1 - p0));
    const int t = threadIdx.x;
    for (int r = t; r < nr; r += 256) {
        const int64_t p = p0 + r / 3;
        const int d = r % 3;
        const bool fr = !cmask[3 * p + d];
        for (int f = 0; f < 3; f++) B[r][f] = (fr && f == d) ? 1.f : 0.f;
        for (int f = 0; f < GKF; f++) B[r][3 + f] = (fr && use_fields) ? O[p * GNO + 3 * f + d] : 0.f;
    }
    __syncthreads();
    if (t < GNB * GNB) {
        const int a = t / GNB, b = t % GNB;
        double s = 0.0;
        for (int r = 0; r < nr; r++) s += (double)B[r][a] * (double)B[r][b];
        G[a][b] = s;
    }
    __syncthreads();
    if (t == 0) {
        double tr = 0.0;
        for (int a = 0; a < GNB; a++) tr += G[a][a];
        const double lam = G_RIDGE * tr / GNB + 1e-30;
        // Cholesky of G + lam I (lower)
        for (int a = 0; a < GNB; a++)
            for (int b = 0; b <= a; b++) {
                double s = G[a][b] + (a == b ? lam : 0.0);
                for (int k = 0; k < b; k++) s -= Lc[a][k] * Lc[b][k];
                Lc[a][b] = (a == b) ? sqrt(fmax(s, 1e-300)) : s / Lc[b][b];
            }
        lam_sh = lam;
        c2w_sh = 0.0;
    }
    __syncthreads();
    const double lam = lam_sh;
    float dBr[G_LSQ_ROWS][GKF];
#pragma unroll
    for (int u = 0; u < G_LSQ_ROWS; u++)
#pragma unroll
        for (int f = 0; f < GKF; f++) dBr[u][f] = 0.f;
    double lacc = 0.0;
    const int qs = t >> 4, tg = t & 15;  // phase 1: sample of the chunk, lane in its group
    for (int q0 = 0; q0 < ns; q0 += G_LSQ_CHUNK) {
        const int nq = min(G_LSQ_CHUNK, ns - q0);
        // 1. B^T e
        if (qs < nq) {
            const int64_t sI = sidx ? sidx[q0 + qs] : q0 + qs;
            const float *e = E + sI * m + 3 * p0;
            float part[GNB];
#pragma unroll
            for (int a = 0; a < GNB; a++) part[a] = 0.f;
            for (int r = tg; r < nr; r += 16) {
                const float ev = e[r];
#pragma unroll
                for (int a = 0; a < GNB; a++) part[a] += B[r][a] * ev;
            }
#pragma unroll
            for (int o = 1; o < 16; o <<= 1)
#pragma unroll
                for (int a = 0; a < GNB; a++) part[a] += __shfl_xor(part[a], o);
            if (tg == 0) {
#pragma unroll
                for (int a = 0; a < GNB; a++) red[qs][a] = part[a];
                srow[qs] = sI * m + 3 * p0;
                wq[qs] = einv[sI] * wscale;
            }
        }
        __syncthreads();
        // 2. c = (G + lam I)^-1 B^T e by the Cholesky factor, one thread per sample
        if (t < nq) {
            double y[GNB], c[GNB];
            for (int a = 0; a < GNB; a++) {
                double s = red[t][a];
                for (int k = 0; k < a; k++) s -= Lc[a][k] * y[k];
                y[a] = s / Lc[a][a];
            }
            for (int a = GNB - 1; a >= 0; a--) {
                double s = y[a];
                for (int k = a + 1; k < GNB; k++) s -= Lc[k][a] * c[k];
                c[a] = s / Lc[a][a];
            }
            double c2 = 0.0;
            for (int a = 0; a < GNB; a++) {
                cv[t][a] = c[a];
                c2 += c[a] * c[a];
            }
            // loss: lambda |c|^2 of the sample; the ridge gradient's weight sum
            lacc += (double)wq[t] * lam * c2;
            atomicAdd(&c2w_sh, (double)wq[t] * c2);
        }
        __syncthreads();
        // 3. residuals of the thread's rows (each row of B read once) for every sample of the
        // chunk
#pragma unroll
        for (int u = 0; u < G_LSQ_ROWS; u++) {
            const int r = t + 256 * u;
            if (r >= nr) break;
            float br[GNB];
#pragma unroll
            for (int a = 0; a < GNB; a++) br[a] = B[r][a];
            for (int q = 0; q < nq; q++) {
                double bc = 0.0;
#pragma unroll
                for (int a = 0; a < GNB; a++) bc += (double)br[a] * cv[q][a];
                const float res = (float)((double)E[srow[q] + r] - bc);
                const float w = wq[q];
                lacc += (double)w * ((double)res * res);
#pragma unroll
                for (int f = 0; f < GKF; f++) dBr[u][f] += -2.f * w * res * (float)cv[q][3 + f];
            }
        }
        __syncthreads();  // red, cv, wq and srow are rewritten by the next chunk
    }
    // lambda = ridge tr(B^T B) / (3 + K) depends on the fields too: d lambda / dB = 2 ridge B / (3 + K)
    const float rg = (float)(2.0 * G_RIDGE / GNB * c2w_sh);
#pragma unroll
    for (int u = 0; u < G_LSQ_ROWS; u++) {
        const int r = t + 256 * u;
        if (r >= nr) break;
        const int64_t p = p0 + r / 3;
        const int d = r % 3;
#pragma unroll
        for (int f = 0; f < GKF; f++) dO[p * GNO + 3 * f + d] = dBr[u][f] + rg * B[r][3 + f];
    }
    lsum[t] = lacc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) lsum[t] += lsum[t + o];
        __syncthreads();
    }
    if (t == 0) lpart[I] = lsum[0];
}

// sum of the loss partials into out[0]
__global__ void k_gnn_lsum(int64_t nagg, const double *lpart, double *out)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < nagg; k += 256) s += lpart[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// ----------------------------------------------------------------------------------------
// backward
// ----------------------------------------------------------------------------------------
// dH = W_d dO (GNO -> 64); one wavefront per point, lane = feature
__global__ __launch_bounds__(256) void k_gnn_decode_bwd(int64_t n, int64_t npad, const float *theta,
                                                        const float *dO, float *dH)
{
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int f = threadIdx.x & 63;
    if (p >= npad) return;
    float s = 0.f;
    if (p < n)
        for (int o = 0; o < GNO; o++) s += theta[OFF_WD + f * GNO + o] * dO[p * GNO + o];
    dH[p * GF + f] = s;
}

// partial outer products sum_p A_p B_p^T and sum_p B_p over chunks of points: output
// [(ka + 1) x kb] per workgroup (row ka = bias gradient).  The chunk's rows of A and B are
// staged in LDS (dynamic: chunk (ka + kb) floats) and every output sums them from there.
template <typename TA>
__global__ __launch_bounds__(256) void k_gnn_outer(int64_t npad, int64_t chunk, const TA *A, int ka,
                                                   const float *B, int kb, float *part)
{
    extern __shared__ float osm[];
    const int64_t p0 = (int64_t)blockIdx.x * chunk;
    const int np_ = (int)(min(npad, p0 + chunk) - p0);
    float *const As = osm;
    float *const Bs = osm + chunk * ka;
    for (int i = threadIdx.x; i < np_ * ka; i += 256) {
        if constexpr (sizeof(TA) == 2) As[i] = bf2f(A[p0 * ka + i]);
        else As[i] = A[p0 * ka + i];
    }
    for (int i = threadIdx.x; i < np_ * kb; i += 256) Bs[i] = B[p0 * kb + i];
    __syncthreads();
    const int tot = (ka + 1) * kb;
    for (int e = threadIdx.x; e < tot; e += 256) {
        const int a = e / kb, b = e - a * kb;
        float s = 0.f;
        if (a < ka)
            for (int p = 0; p < np_; p++) s += As[p * ka + a] * Bs[p * kb + b];
        else
            for (int p = 0; p < np_; p++) s += Bs[p * kb + b];
        part[(int64_t)blockIdx.x * tot + e] = s;
    }
}

// out[e] += sum_w part[w * len + e]: a workgroup sums 16 outputs, each over 16 threads that
// take every 16th partial; the 16 sums are added in a fixed order (the result does not depend
// on scheduling)
__global__ __launch_bounds__(256) void k_gnn_reduce_add(int64_t nw, int64_t len, const float *part, float *out)
{
    __shared__ float sh[16][17];
    const int el = threadIdx.x & 15, ws = threadIdx.x >> 4;
    const int64_t e = (int64_t)blockIdx.x * 16 + el;
    float s = 0.f;
    if (e < len)
        for (int64_t w = ws; w < nw; w += 16) s += part[w * len + e];
    sh[ws][el] = s;
    __syncthreads();
    if (threadIdx.x < 16 && e < len) {
        float t = 0.f;
#pragma unroll
        for (int k = 0; k < 16; k++) t += sh[k][el];
        out[e] += t;
    }
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

// error sample e = z - t on the dofs (t = M_T(J z)), stored in float, and block partials of
// |e|^2 over free dofs
__global__ __launch_bounds__(256) void k_gnn_store_err(int64_t m, const double *z, const double *tt,
                                                       const uint8_t *cmask, float *E, double *part)
{
    __shared__ double sh[256];
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (k < m) {
        const double e = cmask[k] ? 0.0 : z[k] - tt[k];
        E[k] = (float)e;
        s = e * e;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// einv[slot] = 1 / sum(part) (0 for a zero sample)
__global__ void k_gnn_einv(int nblk, const double *part, float *einv, int64_t slot)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < nblk; k += 256) s += part[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) einv[slot] = sh[0] > 0.0 ? (float)(1.0 / sh[0]) : 0.f;
}

// ----------------------------------------------------------------------------------------
// host
// ----------------------------------------------------------------------------------------
void pdx_gnn_free(pdx_ctx *c)
{
    pdx_gnn *g = c->gnn;
    if (!g) return;
    std::vector<void *> ps = {g->near_ptr, g->near, g->theta, g->grad, g->am, g->av, g->F, g->Z0,
                              g->O, g->S1, g->S2, g->Sg, g->lpart, g->sidx, g->dH, g->dH2,
                              g->dZf, g->dO, g->dZ, g->dX, g->wpart, g->E, g->einv, g->tmp,
                              g->tmp2};
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
    return PDX_OK;
}

static pdx_status gnn_wcopy(pdx_ctx *c, pdx_gnn *g)
{
    hipLaunchKernelGGL(k_gnn_wcopy, g1(GL * GK * GF), dim3(256), 0, c->stream, g->theta, g->wT[0],
                       g->wN[0], g->wT[1], g->wN[1], g->wT[2], g->wN[2]);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

// He-style initial weights; the decoder starts small and random (a zero decoder would give
// zero fields, which receive no gradient)
static pdx_status gnn_init_params(pdx_ctx *c, pdx_gnn *g, uint64_t seed)
{
    std::vector<float> th(G_NPARAM, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    for (int k = 0; k < GFIN * GF; k++) th[OFF_WE + k] = nd(rng) * std::sqrt(2.f / GFIN);
    for (int l = 0; l < GL; l++)
        for (int k = 0; k < GK * GF; k++) th[OFF_W(l) + k] = 0.5f * nd(rng) * std::sqrt(2.f / GK);
    for (int k = 0; k < GF * GNO; k++) th[OFF_WD + k] = nd(rng) * std::sqrt(1.f / GF);
    PDX_HIP(c, hipMemcpy(g->theta, th.data(), sizeof(float) * G_NPARAM, hipMemcpyHostToDevice));
    PDX_HIP(c, hipMemset(g->am, 0, sizeof(float) * G_NPARAM));
    PDX_HIP(c, hipMemset(g->av, 0, sizeof(float) * G_NPARAM));
    g->adam_t = 0;
    g->have_fields = false;
    return gnn_wcopy(c, g);
}

// graph, input features and activation buffers (once per context, after the constrained
// dofs are known)
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
    PDX_TRY(pdx_nk_aggregates(c));
    g->agg = c->nk->agg;
    g->nagg = c->nk->nagg;
    if (g->agg > G_AGG_MAX)
        return pdx_fail(c, PDX_ERR_UNSUPPORTED, "the learned preconditioner takes at most %lld points",
                        (long long)G_AGG_MAX * 256);
    const double lmax = G_NEAR_FRAC * c->P.delta;
    int32_t *dcnt = nullptr;
    PDX_TRY(pdx_alloc(c, &dcnt, n));
    hipLaunchKernelGGL(k_near_count, g1(n), dim3(256), 0, c->stream, n, c->slice_ptr, c->cnt, c->L,
                       lmax, dcnt);
    PDX_HIP(c, hipGetLastError());
    std::vector<int32_t> nc(n), bc(n);
    std::vector<uint8_t> cm(3 * n);
    PDX_HIP(c, hipMemcpyAsync(nc.data(), dcnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
    PDX_HIP(c, hipMemcpyAsync(bc.data(), c->cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
    PDX_HIP(c, hipMemcpyAsync(cm.data(), c->nk->cmask, 3 * n, hipMemcpyDeviceToHost, c->stream));
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
    // input features from the reference positions
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
    std::vector<float> F((size_t)g->npad * GFIN, 0.f);
    for (int64_t I = 0; I < g->nagg; I++) {
        const int64_t p0 = I * g->agg, p1 = std::min(n, p0 + g->agg);
        double cen[3] = {0, 0, 0};
        for (int64_t p = p0; p < p1; p++)
            for (int d = 0; d < 3; d++) cen[d] += Xr[4 * p + d];
        for (int d = 0; d < 3; d++) cen[d] /= (double)(p1 - p0);
        double h = 0.0;
        for (int64_t p = p0; p < p1; p++)
            for (int d = 0; d < 3; d++) h = std::max(h, std::fabs(Xr[4 * p + d] - cen[d]));
        h = std::max(h, 1e-300);
        for (int64_t p = p0; p < p1; p++) {
            float *f = &F[p * GFIN];
            for (int d = 0; d < 3; d++) {
                const double half = std::max(0.5 * (hi[d] - lo[d]), 1e-300);
                f[d] = (float)((Xr[4 * p + d] - 0.5 * (hi[d] + lo[d])) / half);
                const double rel = (Xr[4 * p + d] - cen[d]) / h;
                f[3 + d] = (float)rel;
                f[6 + d] = cm[3 * p + d] ? 1.f : 0.f;
                f[11 + d] = (float)(rel * rel);
            }
            f[9] = (float)nc[p] / (float)nmax;
            f[10] = (float)bc[p] / (float)bmax;
            f[14] = 1.f;
            f[15] = 0.f;
        }
    }
    const int64_t np = g->npad;
    PDX_TRY(pdx_alloc(c, &g->F, np * GFIN));
    PDX_HIP(c, hipMemcpy(g->F, F.data(), sizeof(float) * np * GFIN, hipMemcpyHostToDevice));
    // activations
    PDX_TRY(pdx_alloc(c, &g->Z0, np * GF));
    for (int l = 0; l <= GL; l++) PDX_TRY(pdx_alloc(c, &g->H[l], np * GF));
    for (int l = 0; l < GL; l++) {
        PDX_TRY(pdx_alloc(c, &g->X[l], np * GK));
        PDX_TRY(pdx_alloc(c, &g->Z[l], np * GF));
    }
    PDX_TRY(pdx_alloc(c, &g->O, np * GNO));
    PDX_TRY(pdx_alloc(c, &g->S1, g->nb1 * GF * 3));
    PDX_TRY(pdx_alloc(c, &g->S2, g->nb2 * GF * 3));
    PDX_TRY(pdx_alloc(c, &g->Sg, GF * 3));
    PDX_TRY(pdx_alloc(c, &g->lpart, g->nagg + 8));
    PDX_TRY(pdx_alloc(c, &g->dH, np * GF));
    PDX_TRY(pdx_alloc(c, &g->dH2, np * GF));
    PDX_TRY(pdx_alloc(c, &g->dZf, np * GF));
    PDX_TRY(pdx_alloc(c, &g->dO, np * GNO));
    PDX_TRY(pdx_alloc(c, &g->dZ, np * GF));
    PDX_TRY(pdx_alloc(c, &g->dX, np * GK));
    g->wpart_len = std::max<int64_t>((np / GDW_CHUNK) * GK * GF, (np / 256 + 1) * (GK + 1) * GF);
    PDX_TRY(pdx_alloc(c, &g->wpart, g->wpart_len));
    PDX_TRY(pdx_alloc(c, &g->tmp, 3 * n));
    PDX_TRY(pdx_alloc(c, &g->tmp2, (3 * n + 255) / 256 + 8));
    PDX_TRY(gnn_alloc_params(c, g));
    g->ready = true;
    return PDX_OK;
}

// the fields of the current parameters into O; TRAIN keeps the activations for backward
static pdx_status gnn_forward(pdx_ctx *c, bool train)
{
    pdx_gnn *g = c->gnn;
    const int64_t n = g->n, np = g->npad;
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_gnn_encode, dim3((unsigned)(np / 4)), dim3(256), 0, s, n, np, g->F, g->theta,
                       train ? g->Z0 : nullptr, g->H[0]);
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
    const int ppb = 256 / GNO;
    hipLaunchKernelGGL(k_gnn_decode, dim3((unsigned)((np + ppb - 1) / ppb)), dim3(256), 0, s, n, np,
                       g->H[GL], g->theta, g->O);
    PDX_HIP(c, hipGetLastError());
    g->have_fields = !train;
    return PDX_OK;
}

// loss of the samples sidx[0 .. ns) (all of E[0 .. ns) when sidx is null) for the fields in
// O (or the translations alone), and dO; the loss lands in tmp2[0]
static pdx_status gnn_loss(pdx_ctx *c, const int32_t *sidx, int ns, bool use_fields, double *loss)
{
    pdx_gnn *g = c->gnn;
    hipStream_t s = c->stream;
    PDX_HIP(c, hipMemsetAsync(g->dO, 0, sizeof(float) * g->npad * GNO, s));
    hipLaunchKernelGGL(k_gnn_lsq, dim3((unsigned)g->nagg), dim3(256), 0, s, g->n, g->agg, g->O,
                       c->nk->cmask, g->E, g->einv, sidx, ns, 3 * g->n, use_fields ? 1 : 0,
                       1.f / (float)ns, g->dO, g->lpart);
    hipLaunchKernelGGL(k_gnn_lsum, dim3(1), dim3(256), 0, s, g->nagg, g->lpart, g->tmp2);
    PDX_HIP(c, hipGetLastError());
    if (loss) {
        PDX_HIP(c, hipMemcpyAsync(loss, g->tmp2, sizeof(double), hipMemcpyDeviceToHost, s));
        PDX_HIP(c, hipStreamSynchronize(s));
    }
    return PDX_OK;
}

// adds the gradient of the last training forward pass (dO from gnn_loss) to g->grad
static pdx_status gnn_backward(pdx_ctx *c)
{
    pdx_gnn *g = c->gnn;
    const int64_t n = g->n, np = g->npad;
    hipStream_t s = c->stream;
    const int64_t nch = np / 256;  // 256-point chunks
    hipLaunchKernelGGL(k_gnn_decode_bwd, dim3((unsigned)(np / 4)), dim3(256), 0, s, n, np, g->theta, g->dO, g->dH);
    // W_d, b_d
    hipLaunchKernelGGL(k_gnn_outer<uint16_t>, dim3((unsigned)nch), dim3(256), 256 * (GF + GNO) * sizeof(float),
                       s, np, (int64_t)256, g->H[GL], GF, g->dO, GNO, g->wpart);
    hipLaunchKernelGGL(k_gnn_reduce_add, g16((GF + 1) * GNO), dim3(256), 0, s, nch, (int64_t)(GF + 1) * GNO,
                       g->wpart, g->grad + OFF_WD);
    PDX_HIP(c, hipGetLastError());
    float *dHo = g->dH, *dHi = g->dH2;
    for (int l = GL - 1; l >= 0; l--) {
        hipLaunchKernelGGL(k_gnn_relu_bwd, dim3((unsigned)nch), dim3(256), 0, s, n, dHo, g->Z[l], g->dZ,
                           nullptr, g->wpart);
        hipLaunchKernelGGL(k_gnn_reduce_add, g16(GF), dim3(256), 0, s, nch, (int64_t)GF, g->wpart,
                           g->grad + OFF_B(l));
        hipLaunchKernelGGL(k_gnn_gemm_dw, dim3((unsigned)(np / GDW_CHUNK)), dim3(256), 0, s, g->X[l],
                           g->dZ, g->wpart);
        hipLaunchKernelGGL(k_gnn_reduce_add, g16(GK * GF), dim3(256), 0, s, np / GDW_CHUNK,
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
    hipLaunchKernelGGL(k_gnn_outer<float>, dim3((unsigned)nch), dim3(256), 256 * (GFIN + GF) * sizeof(float),
                       s, np, (int64_t)256, g->F, GFIN, g->dZf, GF, g->wpart);
    hipLaunchKernelGGL(k_gnn_reduce_add, g16((GFIN + 1) * GF), dim3(256), 0, s, nch,
                       (int64_t)(GFIN + 1) * GF, g->wpart, g->grad + OFF_WE);
    PDX_HIP(c, hipGetLastError());
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

// the network's fields (GKF per point, 3 floats each, internal order) for the coarse space
pdx_status pdx_gnn_fields(pdx_ctx *c, const float **fields, int *n_fields)
{
    PDX_TRY(pdx_gnn_prepare(c));
    pdx_gnn *g = c->gnn;
    if (!g->have_fields) PDX_TRY(gnn_forward(c, false));
    *fields = g->O;
    *n_fields = GKF;
    return PDX_OK;
}

// the preconditioner with the current coarse space (built by pdx_nk_step with the fields)
pdx_status pdx_gnn_apply(pdx_ctx *c, const double *v, double *z) { return pdx_nk_two_level(c, v, z); }

// Newton pairs carry no error sample of their own: the Krylov pairs are exact
pdx_status pdx_gnn_record(pdx_ctx *c, const double *rhs, const double *du) { return PDX_OK; }

// Krylov pair (J z, z) of a solve preconditioned by the translations' two-level
// preconditioner M_T: the error sample e = z - M_T(J z), reservoir-sampled into cap slots
pdx_status pdx_gnn_record_krylov(pdx_ctx *c, const double *z, const double *jz)
{
    pdx_gnn *g = c->gnn;
    if (!g || !g->collecting || g->cap <= 0) return PDX_OK;
    pdx_nk *k = c->nk;
    if (!k->coarse_ready || k->nf != 3) return PDX_OK;  // collection runs under M_T
    const int64_t m = 3 * c->n;
    if (!g->E) {
        PDX_TRY(pdx_alloc(c, &g->E, (size_t)g->cap * m));
        PDX_TRY(pdx_alloc(c, &g->einv, (size_t)g->cap));
    }
    int64_t slot = -1;
    if (g->seen < g->cap) slot = g->seen;
    else {
        const uint64_t r = g->rng() % (uint64_t)(g->seen + 1);
        if ((int64_t)r < g->cap) slot = (int64_t)r;
    }
    g->seen++;
    if (slot < 0) return PDX_OK;
    PDX_TRY(pdx_nk_two_level(c, jz, g->tmp));
    const int nb = (int)((m + 255) / 256);
    hipLaunchKernelGGL(k_gnn_store_err, dim3(nb), dim3(256), 0, c->stream, m, z, g->tmp, k->cmask,
                       g->E + slot * m, g->tmp2);
    hipLaunchKernelGGL(k_gnn_einv, dim3(1), dim3(256), 0, c->stream, nb, g->tmp2, g->einv, slot);
    PDX_HIP(c, hipGetLastError());
    g->n_s = std::min(g->seen, g->cap);
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

extern "C" pdx_status pdx_gnn_collect(pdx_ctx *c, int32_t enable, int64_t max_samples)
{
    if (!c) return PDX_ERR_ARG;
    if (max_samples < 0) return pdx_fail(c, PDX_ERR_ARG, "max_krylov_pairs < 0");
    PDX_TRY(gnn_ready_for_io(c));
    pdx_gnn *g = c->gnn;
    if (enable && !g->collecting && g->E == nullptr) g->cap = max_samples;
    g->collecting = enable != 0;
    return PDX_OK;
}

extern "C" pdx_status pdx_gnn_pairs(pdx_ctx *c, int64_t *n_newton, int64_t *n_krylov)
{
    if (!c) return PDX_ERR_ARG;
    pdx_gnn *g = c->gnn;
    if (n_newton) *n_newton = 0;
    if (n_krylov) *n_krylov = g ? g->n_s : 0;
    return PDX_OK;
}

static pdx_status gnn_adam_step(pdx_ctx *c, double lr_k)
{
    pdx_gnn *g = c->gnn;
    const float b1 = 0.9f, b2 = 0.999f;
    g->adam_t++;
    const float c1 = 1.f - std::pow(b1, (float)g->adam_t), c2 = 1.f - std::pow(b2, (float)g->adam_t);
    hipLaunchKernelGGL(k_gnn_adam, g1(G_NPARAM), dim3(256), 0, c->stream, (int64_t)G_NPARAM, g->theta,
                       g->grad, g->am, g->av, (float)lr_k, b1, b2, c1, c2, 1e-8f);
    PDX_HIP(c, hipGetLastError());
    return gnn_wcopy(c, g);
}

extern "C" pdx_status pdx_gnn_train(pdx_ctx *c, int32_t steps, int32_t batch, double lr,
                                    uint64_t seed, double *loss_out)
{
    if (!c) return PDX_ERR_ARG;
    if (steps < 0 || batch <= 0 || !(lr > 0)) return pdx_fail(c, PDX_ERR_ARG, "invalid training options");
    PDX_TRY(gnn_ready_for_io(c));
    pdx_gnn *g = c->gnn;
    if (g->n_s == 0) return pdx_fail(c, PDX_ERR_STATE, "no training samples (pdx_gnn_collect)");
    std::mt19937_64 rng(seed);
    // hold out the last tenth of the samples (if there are enough) for validation
    g->n_val = g->n_s >= 20 ? g->n_s / 10 : 0;
    const int64_t ntr = g->n_s - g->n_val;
    const int nb = (int)std::min<int64_t>(batch, ntr);
    if (!g->sidx) PDX_TRY(pdx_alloc(c, &g->sidx, (size_t)std::max<int64_t>(g->n_s, 1)));
    if (!g->trained) PDX_TRY(gnn_init_params(c, g, seed));
    std::vector<int32_t> order(ntr);
    for (int64_t i = 0; i < ntr; i++) order[i] = (int32_t)i;
    double last = 0.0;
    for (int it = 0; it < steps; it++) {
        // a random subset of nb training samples
        for (int q = 0; q < nb; q++) std::swap(order[q], order[q + rng() % (uint64_t)(ntr - q)]);
        PDX_HIP(c, hipMemcpyAsync(g->sidx, order.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice,
                                  c->stream));
        PDX_HIP(c, hipMemsetAsync(g->grad, 0, sizeof(float) * G_NPARAM, c->stream));
        PDX_TRY(gnn_forward(c, true));
        const bool report = it == steps - 1 || (it % 25) == 24;
        PDX_TRY(gnn_loss(c, g->sidx, nb, true, report ? &last : nullptr));
        PDX_TRY(gnn_backward(c));
        // cosine decay of the step size
        PDX_TRY(gnn_adam_step(c, lr * 0.5 * (1.0 + std::cos(M_PI * it / std::max(steps, 1)))));
    }
    PDX_TRY(gnn_forward(c, false));
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    g->trained = true;
    if (loss_out) *loss_out = last;
    // held-out loss with the fields over the loss of the translations alone (1 = no gain)
    g->val_loss = 1.0;
    if (g->n_val > 0) {
        std::vector<int32_t> vi(g->n_val);
        for (int64_t q = 0; q < g->n_val; q++) vi[q] = (int32_t)(ntr + q);
        PDX_HIP(c, hipMemcpyAsync(g->sidx, vi.data(), sizeof(int32_t) * g->n_val, hipMemcpyHostToDevice,
                                  c->stream));
        double lf = 0.0, lt = 0.0;
        PDX_TRY(gnn_loss(c, g->sidx, (int)g->n_val, true, &lf));
        PDX_TRY(gnn_loss(c, g->sidx, (int)g->n_val, false, &lt));
        g->val_loss_t = lt;
        g->val_loss = lt > 0 ? lf / lt : 1.0;
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
        g->trained = true;
        g->have_fields = false;
    }
    if (out_scale) *out_scale = 1.0;
    if (theta_out) PDX_HIP(c, hipMemcpy(theta_out, g->theta, sizeof(float) * G_NPARAM, hipMemcpyDeviceToHost));
    PDX_HIP(c, hipDeviceSynchronize());
    return PDX_OK;
}

extern "C" int64_t pdx_gnn_n_params(void) { return G_NPARAM; }

extern "C" pdx_status pdx_gnn_graph(pdx_ctx *c, int64_t *n_near, int64_t *perm, int64_t *near_ptr,
                                    int32_t *near, float *feat, double *dinv, uint8_t *cmask)
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
    if (feat) PDX_HIP(c, hipMemcpy(feat, g->F, sizeof(float) * GFIN * n, hipMemcpyDeviceToHost));
    if (dinv) {
        PDX_TRY(pdx_nk_build_jacobi(c));
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
    const float *fl = nullptr;
    int nl = 0;
    PDX_TRY(pdx_gnn_fields(c, &fl, &nl));
    PDX_TRY(pdx_nk_build_jacobi(c));
    PDX_TRY(pdx_nk_build_coarse_fields(c, fl, nl));
    PDX_HIP(c, hipMemcpy(g->tmp, y, sizeof(double) * m, hipMemcpyHostToDevice));
    double *zd = c->nk->w;
    PDX_TRY(pdx_nk_two_level(c, g->tmp, zd));
    PDX_HIP(c, hipMemcpy(z, zd, sizeof(double) * m, hipMemcpyDeviceToHost));
    return PDX_OK;
}

extern "C" pdx_status pdx_gnn_field_values(pdx_ctx *c, float *fields)
{
    if (!c || !fields) return PDX_ERR_ARG;
    PDX_TRY(gnn_ready_for_io(c));
    pdx_gnn *g = c->gnn;
    if (!g->trained) return pdx_fail(c, PDX_ERR_STATE, "the network has no parameters (pdx_gnn_params)");
    PDX_TRY(gnn_forward(c, false));
    PDX_HIP(c, hipMemcpy(fields, g->O, sizeof(float) * GNO * c->n, hipMemcpyDeviceToHost));
    return PDX_OK;
}

extern "C" pdx_status pdx_gnn_loss_grad(pdx_ctx *c, int32_t n_samples, const double *Y,
                                        const double *X, double *loss, float *grad)
{
    if (!c || n_samples <= 0 || !Y) return PDX_ERR_ARG;
    PDX_TRY(gnn_ready_for_io(c));
    pdx_gnn *g = c->gnn;
    if (!g->trained) return pdx_fail(c, PDX_ERR_STATE, "the network has no parameters (pdx_gnn_params)");
    const int64_t m = 3 * c->n;
    // the samples replace the collected ones
    if (g->E && g->cap < n_samples) {
        (void)hipFree(g->E);
        (void)hipFree(g->einv);
        g->E = nullptr;
        g->einv = nullptr;
    }
    if (!g->E) {
        g->cap = std::max<int64_t>(g->cap, n_samples);
        PDX_TRY(pdx_alloc(c, &g->E, (size_t)g->cap * m));
        PDX_TRY(pdx_alloc(c, &g->einv, (size_t)g->cap));
    }
    const int nb = (int)((m + 255) / 256);
    double *zd = c->nk->w;
    PDX_HIP(c, hipMemsetAsync(g->tmp, 0, sizeof(double) * m, c->stream));
    for (int s = 0; s < n_samples; s++) {
        PDX_HIP(c, hipMemcpyAsync(zd, Y + s * m, sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_gnn_store_err, dim3(nb), dim3(256), 0, c->stream, m, zd, g->tmp,
                           c->nk->cmask, g->E + s * m, g->tmp2);
        hipLaunchKernelGGL(k_gnn_einv, dim3(1), dim3(256), 0, c->stream, nb, g->tmp2, g->einv, (int64_t)s);
        PDX_HIP(c, hipStreamSynchronize(c->stream));
    }
    g->n_s = g->seen = n_samples;
    PDX_HIP(c, hipMemsetAsync(g->grad, 0, sizeof(float) * G_NPARAM, c->stream));
    PDX_TRY(gnn_forward(c, true));
    double l = 0.0;
    PDX_TRY(gnn_loss(c, nullptr, n_samples, true, &l));
    PDX_TRY(gnn_backward(c));
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = l;
    if (grad) PDX_HIP(c, hipMemcpy(grad, g->grad, sizeof(float) * G_NPARAM, hipMemcpyDeviceToHost));
    return PDX_OK;
}
