// Learned warm start: a 24 -> 128 -> 128 -> 3 per-vertex MLP (SiLU) whose
// GEMMs run on gfx950 BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation, trained on the device with Adam, and the learned solver step.
//
// MFMA 16x16x32 operand layout (wave64, lane l):
//   A (16 x 32, M x K): row l % 16, k = 8 * (l / 16) + 0..7
//   B (32 x 16, K x N): col l % 16, k = 8 * (l / 16) + 0..7
//   D (16 x 16):        col l % 16, rows 4 * (l / 16) + 0..3
//
// The learned step (no counterpart in the reference): after the inertial
// prediction x~, the network predicts the solver's correction c; the solve
// starts at x0 = x~ + c with the multipliers that satisfy the constraint
// equations there, lambda0 = -C(x0) / alpha~, and runs the configured number
// of Gauss-Seidel / Jacobi iterations (which resolve contacts and whatever the
// network did not predict).  If the constraint residual then exceeds
// ratio * reference, the step is redone from the saved state with the plain
// cold-start solve (the reference solver's semantics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "capi_internal.h"
#include "kernels.h"
#include "xpbd_math.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kIn = XPBD_MLP_IN, kInP = 32, kH = XPBD_MLP_HIDDEN, kOut = XPBD_MLP_OUT, kOutP = 16;
// parameter offsets in the FP32 master vector
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kOut * kH, kNParams = oB3 + kOut;
static_assert(kNParams == 20099, "parameter count");

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
// inference: hardware reciprocal (1 ulp) in place of the IEEE division; the
// result is rounded to BF16 right after
__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// ---------------------------------------------------------------------------
// Generic GEMM  C[m, n] = alpha * sum_k A(m, k) B(n, k) + beta * C[m, n]
// A(m, k) = A[m * am + k * ak], B(n, k) = B[n * bm + k * bk] (FP32 in memory,
// rounded to BF16 while staged into LDS), C row-major with ldc.  Workgroup:
// 4 waves, 32 x 32 output tile (2 x 2 MFMA tiles), K step 32.
// ---------------------------------------------------------------------------
constexpr int kGT = 32;
constexpr int kMaxSplit = 64;

// blockIdx.z selects a K range of length kchunk; with nsplit > 1 the partial
// products go to Cpart[z][M][N] (reduced by k_reduce_split), else to C.
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, int kchunk, float alpha,
                                             const float* __restrict__ A, int am, int ak,
                                             const float* __restrict__ B, int bm, int bk, float beta,
                                             float* __restrict__ Cm, int ldc, float* __restrict__ Cpart)
{
    __shared__ alignas(16) __bf16 sA[kGT][kGT + 8];  // [m][k], padded rows keep 16-byte alignment
    __shared__ alignas(16) __bf16 sB[kGT][kGT + 8];  // [n][k]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    const int tm = (wv >> 1) * 16, tn = (wv & 1) * 16;  // this wave's 16 x 16 sub-tile
    // staging order: along k when k is the unit-stride index, else along m / n
    const bool a_kfast = ak == 1, b_kfast = bk == 1;
    f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
    for (int k0 = kb; k0 < ke; k0 += kGT) {
        for (int i = tid; i < kGT * kGT; i += 256) {
            const int r = a_kfast ? i / kGT : i % kGT, k = a_kfast ? i % kGT : i / kGT;
            const int gm = m0 + r, gk = k0 + k;
            sA[r][k] = (__bf16)((gm < M && gk < ke) ? A[(size_t)gm * am + (size_t)gk * ak] : 0.0f);
        }
        for (int i = tid; i < kGT * kGT; i += 256) {
            const int r = b_kfast ? i / kGT : i % kGT, k = b_kfast ? i % kGT : i / kGT;
            const int gn = n0 + r, gk = k0 + k;
            sB[r][k] = (__bf16)((gn < N && gk < ke) ? B[(size_t)gn * bm + (size_t)gk * bk] : 0.0f);
        }
        __syncthreads();
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[tm + (lane & 15)][8 * (lane >> 4)]);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[tn + (lane & 15)][8 * (lane >> 4)]);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        __syncthreads();
    }
    const int col = n0 + tn + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int row = m0 + tm + 4 * (lane >> 4) + j;
        if (row < M && col < N) {
            if (Cpart) {
                Cpart[((size_t)blockIdx.z * M + row) * N + col] = acc[j];
            } else {
                float* c = Cm + (size_t)row * ldc + col;
                *c = alpha * acc[j] + (beta != 0.f ? beta * *c : 0.f);
            }
        }
    }
}

__global__ void k_reduce_split(const float* __restrict__ part, int nsplit, int M, int N, float alpha,
                               float* __restrict__ Cm, int ldc)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= M * N) return;
    float s = 0.f;
    for (int z = 0; z < nsplit; z++) s += part[(size_t)z * M * N + i];
    Cm[(size_t)(i / N) * ldc + i % N] = alpha * s;
}

// C = alpha A B^T (+ beta C without split).  Small-output products with a long
// K (weight gradients: K = rows of the minibatch) are split over K into the
// workspace `split` (kMaxSplit * M * N floats) and reduced in a fixed order.
void gemm(hipStream_t s, int M, int N, int K, float alpha, const float* A, int am, int ak, const float* B, int bm,
          int bk, float beta, float* C, int ldc, float* split)
{
    const int gx = (N + kGT - 1) / kGT, gy = (M + kGT - 1) / kGT;
    int nsplit = 1;
    if (split && beta == 0.f && gx * gy < 256 && K >= 8 * kGT) {
        nsplit = (512 + gx * gy - 1) / (gx * gy);
        const int maxs = (K + 4 * kGT - 1) / (4 * kGT);  // at least 4 K steps per split
        if (nsplit > maxs) nsplit = maxs;
        if (nsplit > kMaxSplit) nsplit = kMaxSplit;
    }
    int kchunk = (K + nsplit - 1) / nsplit;
    kchunk = (kchunk + kGT - 1) / kGT * kGT;
    nsplit = (K + kchunk - 1) / kchunk;
    if (nsplit < 1) nsplit = 1;
    const dim3 g(gx, gy, nsplit);
    hipLaunchKernelGGL(k_gemm, g, dim3(256), 0, s, M, N, K, kchunk, alpha, A, am, ak, B, bm, bk, beta, C, ldc,
                       nsplit > 1 ? split : nullptr);
    if (nsplit > 1)
        hipLaunchKernelGGL(k_reduce_split, dim3((M * N + 255) / 256), dim3(256), 0, s, split, nsplit, M, N, alpha, C,
                           ldc);
}

// ---------------------------------------------------------------------------
// elementwise helpers for training
// ---------------------------------------------------------------------------
// X[r][k] standardised into Xs (ld kInP, zero padded)
__global__ void k_standardize(const float* __restrict__ X, int n, const float* __restrict__ mean,
                              const float* __restrict__ istd, float* __restrict__ Xs)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n * kInP) return;
    const int r = i / kInP, k = i % kInP;
    Xs[i] = k < kIn ? (X[(size_t)r * kIn + k] - mean[k]) * istd[k] : 0.f;
}

// Z += bias (per column), H = silu(Z)
__global__ void k_bias_silu(float* __restrict__ Z, const float* __restrict__ b, int n, int ncol,
                            float* __restrict__ Hout)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n * ncol) return;
    const float z = Z[i] + b[i % ncol];
    Z[i] = z;
    if (Hout) Hout[i] = silu(z);
}

__global__ void k_bias(float* __restrict__ Z, const float* __restrict__ b, int n, int ncol)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n * ncol) return;
    Z[i] += b[i % ncol];
}

// Loss in output units of the network: mean over n * 3 of (Y - T / scale)^2,
// i.e. the squared error of the correction measured in units of scale (the
// physical corrections are ~1e-5 m; a loss in metres would leave gradients far
// below Adam's epsilon).  dY = 2 (Y - T / scale) / (n * 3); loss partial sums.
__global__ void k_mse_grad(const float* __restrict__ Y, const float* __restrict__ T, int n, float inv_scale,
                           float* __restrict__ dY, double* __restrict__ lpart)
{
    __shared__ double sh[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double l = 0.0;
    if (i < n * kOut) {
        const float d = Y[i] - T[i] * inv_scale;
        dY[i] = 2.0f * d / (float)(n * kOut);
        l = (double)d * (double)d;
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = sh[0];
}

// dZ = dH * silu'(Z)
__global__ void k_dsilu(float* __restrict__ dH, const float* __restrict__ Z, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dH[i] *= dsilu(Z[i]);
}

// column sums of D[n][ncol] into out[ncol] (bias gradients), one block per column
__global__ void k_colsum(const float* __restrict__ D, int n, int ncol, float* __restrict__ out)
{
    __shared__ float sh[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < n; r += 256) s += D[(size_t)r * ncol + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float c1, float c2, float eps)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// The hidden feature at MFMA K position pos (0..31) of K block m in the
// inference kernel: the order in which a 16 x 16 output tile pair leaves the
// features in a lane (position 8q + i holds 4q + i of the block's first tile for
// i < 4, 4q + i - 4 of its second tile for i >= 4).
__host__ __device__ constexpr int mlp_k_feature(int m, int pos)
{
    return 32 * m + ((pos & 7) < 4 ? 4 * (pos >> 3) + (pos & 7) : 16 + 4 * (pos >> 3) + (pos & 7) - 4);
}

constexpr bool mlp_k_order_is_a_permutation()
{
    for (int m = 0; m < kH / 32; m++)
        for (int f = 32 * m; f < 32 * m + 32; f++) {
            int hits = 0;
            for (int pos = 0; pos < 32; pos++) hits += mlp_k_feature(m, pos) == f;
            if (hits != 1) return false;
        }
    return true;
}
static_assert(mlp_k_order_is_a_permutation(), "every hidden feature at exactly one K position");
// the inference kernel's shapes: one 32-wide K step for the inputs (8 per lane,
// lanes q < 3 full), 16-wide output tiles in pairs, at most 4 outputs (lanes q = 0)
static_assert(kInP == 32 && kIn % 8 == 0 && kIn <= kInP && kH % 32 == 0 && kOut <= 4 && kOutP == 16,
              "shapes of k_mlp_fused");

// FP32 master -> BF16 MFMA layouts of the inference kernel: W1 [128][32] (k >= 24
// zero), W2 [128][128] and W3 [16][128] (rows >= 3 zero) with their hidden K index
// in mlp_k_feature order
__global__ void k_pack(const float* __restrict__ p, __bf16* __restrict__ w1, __bf16* __restrict__ w2,
                       __bf16* __restrict__ w3)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < kH * kInP) {
        const int o = i / kInP, k = i % kInP;
        w1[i] = (__bf16)(k < kIn ? p[oW1 + o * kIn + k] : 0.f);
    }
    if (i < kH * kH) {
        const int o = i / kH, k = i % kH;
        w2[i] = (__bf16)p[oW2 + o * kH + mlp_k_feature(k / 32, k % 32)];
    }
    if (i < kOutP * kH) {
        const int o = i / kH, k = i % kH;
        w3[i] = (__bf16)(o < kOut ? p[oW3 + o * kH + mlp_k_feature(k / 32, k % 32)] : 0.f);
    }
}

// ---------------------------------------------------------------------------
// Fused inference, 4 waves per workgroup, 16 rows per wave per tile.  Every
// layer forms Z^T = W H^T: the weights (BF16, staged once per workgroup in LDS)
// are the MFMA's A operand and the activations its B operand.  A 16 x 16 output
// tile leaves lane (q = lane / 16, c = lane % 16) holding features 4q..4q+3 of
// row c, so the two tiles of a 32-wide K block hold, per lane, features
// 4q + {0..3} and 16 + 4q + {0..3} of the block: after bias and SiLU, rounded to
// BF16, they are the next layer's B operand as they lie, with that layer's
// weights packed in the same K order (mlp_k_feature).  The activations stay in
// registers from input to output; the first layer's B operand is read from the
// input rows directly (lane (q, c): features 8q..8q+7 of row c, zero past 24).
// out[r] (float4) = scale * MLP((x_r - mean) * istd) (* mask[r]), w = 0.
// ---------------------------------------------------------------------------
constexpr int kFwdRows = 64;

__global__ __launch_bounds__(256) void k_mlp_fused(const float* __restrict__ X, int n, const __bf16* __restrict__ w1,
                                                  const __bf16* __restrict__ w2, const __bf16* __restrict__ w3,
                                                  const float* __restrict__ params, const float* __restrict__ mean,
                                                  const float* __restrict__ istd, float scale,
                                                  float4* __restrict__ out, const float* __restrict__ mask)
{
    __shared__ alignas(16) __bf16 sw1[kH][kInP + 8];
    __shared__ alignas(16) __bf16 sw2[kH][kH + 8];
    __shared__ alignas(16) __bf16 sw3[kOutP][kH + 8];
    __shared__ alignas(16) float sb[2 * kH + kOutP];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // weights in 16-byte pieces (8 BF16 values; rows are multiples of 8 long and
    // the padded LDS rows start on 16-byte boundaries)
    for (int i = tid; i < kH * kInP / 8; i += 256)
        *reinterpret_cast<bf16x8*>(&sw1[i / (kInP / 8)][8 * (i % (kInP / 8))]) = reinterpret_cast<const bf16x8*>(w1)[i];
    for (int i = tid; i < kH * kH / 8; i += 256)
        *reinterpret_cast<bf16x8*>(&sw2[i / (kH / 8)][8 * (i % (kH / 8))]) = reinterpret_cast<const bf16x8*>(w2)[i];
    for (int i = tid; i < kOutP * kH / 8; i += 256)
        *reinterpret_cast<bf16x8*>(&sw3[i / (kH / 8)][8 * (i % (kH / 8))]) = reinterpret_cast<const bf16x8*>(w3)[i];
    for (int i = tid; i < kH; i += 256) {
        sb[i] = params[oB1 + i];
        sb[kH + i] = params[oB2 + i];
    }
    if (tid < kOutP) sb[2 * kH + tid] = tid < kOut ? params[oB3 + tid] : 0.f;
    __syncthreads();
    // from here on the waves share only the weights and biases, which they read
    const int c = lane & 15, q = lane >> 4, kq = 8 * q;
    const bool xin = kq < kIn;   // lanes q < 3 hold input features, q = 3 the zero padding
    float xm[8], xs[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        xm[i] = xin ? mean[kq + i] : 0.f;
        xs[i] = xin ? istd[kq + i] : 0.f;
    }
    const int ntile = (n + kFwdRows - 1) / kFwdRows;
    // the next tile's input features are loaded while the current tile is computed
    float4 xa = make_float4(0.f, 0.f, 0.f, 0.f), xb = xa;
    auto load_row = [&](int t) {
        const int r = t * kFwdRows + wv * 16 + c;
        xa = xb = make_float4(0.f, 0.f, 0.f, 0.f);
        if (xin && r < n) {
            const float4* x4 = reinterpret_cast<const float4*>(X + (size_t)r * kIn + kq);
            xa = x4[0];
            xb = x4[1];
        }
    };
    if ((int)blockIdx.x < ntile) load_row(blockIdx.x);
    for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
        const int r = t * kFwdRows + wv * 16 + c;
        const float xv[8] = { xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w };
        bf16x8 h0;
#pragma unroll
        for (int i = 0; i < 8; i++) h0[i] = (__bf16)((xv[i] - xm[i]) * xs[i]);
        if (t + (int)gridDim.x < ntile) load_row(t + gridDim.x);
        // layer 1: 8 output tiles, one K step; K block m of layer 2 from tiles 2m, 2m + 1
        bf16x8 h1[kH / 32];
#pragma unroll
        for (int m = 0; m < kH / 32; m++) {
            const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(&sw1[32 * m + c][kq]);
            const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(&sw1[32 * m + 16 + c][kq]);
            const f32x4 z0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, h0, f32x4{ 0.f, 0.f, 0.f, 0.f }, 0, 0, 0);
            const f32x4 z1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, h0, f32x4{ 0.f, 0.f, 0.f, 0.f }, 0, 0, 0);
            const float4 b0 = *reinterpret_cast<const float4*>(&sb[32 * m + 4 * q]);
            const float4 b1 = *reinterpret_cast<const float4*>(&sb[32 * m + 16 + 4 * q]);
            h1[m] = bf16x8{ (__bf16)silu_fast(z0[0] + b0.x), (__bf16)silu_fast(z0[1] + b0.y),
                            (__bf16)silu_fast(z0[2] + b0.z), (__bf16)silu_fast(z0[3] + b0.w),
                            (__bf16)silu_fast(z1[0] + b1.x), (__bf16)silu_fast(z1[1] + b1.y),
                            (__bf16)silu_fast(z1[2] + b1.z), (__bf16)silu_fast(z1[3] + b1.w) };
        }
        // layer 2: 8 output tiles x 4 K steps
        f32x4 z2[kH / 16];
#pragma unroll
        for (int nt = 0; nt < kH / 16; nt++) z2[nt] = f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int m = 0; m < kH / 32; m++)
#pragma unroll
            for (int nt = 0; nt < kH / 16; nt++) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sw2[nt * 16 + c][32 * m + kq]);
                z2[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, h1[m], z2[nt], 0, 0, 0);
            }
        bf16x8 h2[kH / 32];
#pragma unroll
        for (int m = 0; m < kH / 32; m++) {
            const float4 b0 = *reinterpret_cast<const float4*>(&sb[kH + 32 * m + 4 * q]);
            const float4 b1 = *reinterpret_cast<const float4*>(&sb[kH + 32 * m + 16 + 4 * q]);
            const f32x4 z0 = z2[2 * m], z1 = z2[2 * m + 1];
            h2[m] = bf16x8{ (__bf16)silu_fast(z0[0] + b0.x), (__bf16)silu_fast(z0[1] + b0.y),
                            (__bf16)silu_fast(z0[2] + b0.z), (__bf16)silu_fast(z0[3] + b0.w),
                            (__bf16)silu_fast(z1[0] + b1.x), (__bf16)silu_fast(z1[1] + b1.y),
                            (__bf16)silu_fast(z1[2] + b1.z), (__bf16)silu_fast(z1[3] + b1.w) };
        }
        // layer 3: one padded output tile x 4 K steps; lanes q = 0 hold outputs 0..3 of row c
        f32x4 y = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int m = 0; m < kH / 32; m++) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sw3[c][32 * m + kq]);
            y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, h2[m], y, 0, 0, 0);
        }
        if (q == 0 && r < n) {
            const float mk = mask ? mask[r] : 1.f;
            out[r] = make_float4(scale * (y[0] + sb[2 * kH]) * mk, scale * (y[1] + sb[2 * kH + 1]) * mk,
                                 scale * (y[2] + sb[2 * kH + 2]) * mk, 0.f);
        }
    }
}

inline unsigned nb(long long n) { return (unsigned)((n + 255) / 256); }

__global__ void k_gather(const float* __restrict__ Xall, const float* __restrict__ Yall,
                         const int* __restrict__ perm, int b0, int n, int rows, float* __restrict__ X,
                         float* __restrict__ T)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= rows * kIn) return;
    const int r = i / kIn, k = i % kIn;
    const int src = perm[(b0 + r) % n];
    X[i] = Xall[(size_t)src * kIn + k];
    if (k < kOut) T[r * kOut + k] = Yall[(size_t)src * kOut + k];
}

// ---------------------------------------------------------------------------
// solver-side kernels of the learned step
// ---------------------------------------------------------------------------
// Per-vertex features after the inertial prediction (row of 24 floats):
//   0-2   inertial displacement of this step (x~ - x)      3-5   velocity
//   6-8   1-ring mean of (u_j - u_i), u = x - X            9-11  1-ring mean of inertial displacement minus own
//   12-14 displacement from rest u_i                       15-17 previous step's solver correction c_i
//   18-20 1-ring mean of (c_j - c_i)                       21    mean log10 E of incident elements
//   22    mean nu of incident elements                     23    fraction of 1-ring neighbours pinned
__global__ void k_features(const float4* __restrict__ pos, const float4* __restrict__ vel,
                           const float4* __restrict__ cprev, const int* __restrict__ nb_off,
                           const int* __restrict__ nbr, const float2* __restrict__ vmat,
                           const float* __restrict__ pinfrac, int nv, float* __restrict__ feat)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nv) return;
    const float4 ui = pos[2 * (size_t)i], pi = pos[2 * (size_t)i + 1], vi = vel[i], ci = cprev[i];
    float su[3] = { 0.f, 0.f, 0.f }, sp[3] = { 0.f, 0.f, 0.f }, sc[3] = { 0.f, 0.f, 0.f };
    const int b = nb_off[i], e = nb_off[i + 1];
    for (int k = b; k < e; k++) {
        const int j = nbr[k];
        const float4 uj = pos[2 * (size_t)j], pj = pos[2 * (size_t)j + 1], cj = cprev[j];
        su[0] += uj.x; su[1] += uj.y; su[2] += uj.z;
        sp[0] += pj.x; sp[1] += pj.y; sp[2] += pj.z;
        sc[0] += cj.x; sc[1] += cj.y; sc[2] += cj.z;
    }
    const float inv = e > b ? 1.0f / (float)(e - b) : 0.f;
    const float own = e > b ? 1.f : 0.f;
    float* f = feat + (size_t)i * kIn;
    f[0] = pi.x; f[1] = pi.y; f[2] = pi.z;
    f[3] = vi.x; f[4] = vi.y; f[5] = vi.z;
    f[6] = su[0] * inv - own * ui.x; f[7] = su[1] * inv - own * ui.y; f[8] = su[2] * inv - own * ui.z;
    f[9] = sp[0] * inv - own * pi.x; f[10] = sp[1] * inv - own * pi.y; f[11] = sp[2] * inv - own * pi.z;
    f[12] = ui.x; f[13] = ui.y; f[14] = ui.z;
    f[15] = ci.x; f[16] = ci.y; f[17] = ci.z;
    f[18] = sc[0] * inv - own * ci.x; f[19] = sc[1] * inv - own * ci.y; f[20] = sc[2] * inv - own * ci.z;
    const float2 mt = vmat[i];
    f[21] = mt.x; f[22] = mt.y; f[23] = pinfrac[i];
}

// in-step motion after the prediction -> pred
__global__ void k_copy_delta(const float4* __restrict__ pos, float4* __restrict__ pred, int nv)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < nv) pred[i] = pos[2 * (size_t)i + 1];
}

// start of the solve moved by the network's correction (free vertices only)
__global__ void k_apply_corr(float4* __restrict__ pos, const float4* __restrict__ vel, const float4* __restrict__ corr,
                             int nv)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nv || vel[i].w != 0.f) return;
    float4 d = pos[2 * (size_t)i + 1];
    const float4 c = corr[i];
    d.x += c.x;
    d.y += c.y;
    d.z += c.z;
    pos[2 * (size_t)i + 1] = d;
}

// solver correction of this step (end of solve minus prediction); 0 at pins
__global__ void k_store_cprev(const float4* __restrict__ pos, const float4* __restrict__ vel,
                              const float4* __restrict__ pred, float4* __restrict__ cprev, int nv)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nv) return;
    const float4 d = pos[2 * (size_t)i + 1], p = pred[i];
    cprev[i] = vel[i].w != 0.f ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(d.x - p.x, d.y - p.y, d.z - p.z, 0.f);
}

__global__ void k_zero_delta(float4* __restrict__ pos, int nv)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < nv) pos[2 * (size_t)i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
}

}  // namespace

// ---------------------------------------------------------------------------
// network object
// ---------------------------------------------------------------------------
static uint64_t next_mlp_version()
{
    static std::atomic<uint64_t> counter{ 0 };
    return ++counter;
}

struct xpbd_mlp {
    int device = 0;           // device current at creation (weights and stream live there)
    hipStream_t stream = nullptr;
    float* d_p = nullptr;     // FP32 master parameters
    float* d_g = nullptr;     // gradient
    float* d_m = nullptr;     // Adam first moment
    float* d_v = nullptr;     // Adam second moment
    __bf16* d_w1 = nullptr;   // BF16 copies in the fused-inference layout
    __bf16* d_w2 = nullptr;
    __bf16* d_w3 = nullptr;
    float* d_mean = nullptr;
    float* d_istd = nullptr;
    float* d_split = nullptr; // split-K partial products
    float scale = 1.f;
    // Identifies the network and its output scale (the scale is a launch
    // argument of the inference, so graphs that captured it go stale when it
    // changes); never reused within the process.
    uint64_t version = 0;
    float h_mean[24] = {}, h_std[24] = {};  // host copies of the scaling (for xpbd_mlp_save)
    bool has_scaling = false;
    int64_t adam_t = 0;
    // training workspace (rows = capacity)
    int cap = 0;
    float *d_X = nullptr, *d_T = nullptr, *d_Xs = nullptr, *d_Z1 = nullptr, *d_H1 = nullptr, *d_Z2 = nullptr,
          *d_H2 = nullptr, *d_Y = nullptr, *d_dY = nullptr, *d_dH = nullptr, *d_dH1 = nullptr;
    double* d_lpart = nullptr;
    float4* d_out = nullptr;
    std::string err;
};

namespace {

int merr(xpbd_mlp* m, int code, const std::string& s)
{
    if (m) m->err = s;
    return code;
}

hipError_t pack(xpbd_mlp* m)
{
    hipLaunchKernelGGL(k_pack, dim3(nb(kH * kH)), dim3(256), 0, m->stream, m->d_p, m->d_w1, m->d_w2, m->d_w3);
    return hipGetLastError();
}

void free_ws(xpbd_mlp* m)
{
    float** bufs[] = { &m->d_X, &m->d_T, &m->d_Xs, &m->d_Z1, &m->d_H1, &m->d_Z2, &m->d_H2, &m->d_Y, &m->d_dY,
                       &m->d_dH, &m->d_dH1 };
    for (float** b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    if (m->d_lpart) (void)hipFree(m->d_lpart);
    if (m->d_out) (void)hipFree(m->d_out);
    m->d_lpart = nullptr;
    m->d_out = nullptr;
    m->cap = 0;
}

hipError_t ensure_ws(xpbd_mlp* m, int rows)
{
    if (rows <= m->cap) return hipSuccess;
    (void)hipStreamSynchronize(m->stream);
    free_ws(m);
    const size_t r = (size_t)rows;
    hipError_t e = hipSuccess;
    auto A = [&](float** p, size_t n) {
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(float));
    };
    A(&m->d_X, r * kIn);
    A(&m->d_T, r * kOut);
    A(&m->d_Xs, r * kInP);
    A(&m->d_Z1, r * kH);
    A(&m->d_H1, r * kH);
    A(&m->d_Z2, r * kH);
    A(&m->d_H2, r * kH);
    A(&m->d_Y, r * kOut);
    A(&m->d_dY, r * kOut);
    A(&m->d_dH, r * kH);
    A(&m->d_dH1, r * kH);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&m->d_lpart), sizeof(double) * nb(r * kOut));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&m->d_out), sizeof(float4) * r);
    if (e == hipSuccess) m->cap = rows;
    else free_ws(m);
    return e;
}

// Forward on the n rows in d_X (n <= cap), keeping the activations for backward.
void forward_train(xpbd_mlp* m, int n)
{
    hipStream_t s = m->stream;
    const float* P = m->d_p;
    hipLaunchKernelGGL(k_standardize, dim3(nb((long long)n * kInP)), dim3(256), 0, s, m->d_X, n, m->d_mean, m->d_istd,
                       m->d_Xs);
    gemm(s, n, kH, kIn, 1.f, m->d_Xs, kInP, 1, P + oW1, kIn, 1, 0.f, m->d_Z1, kH, nullptr);
    hipLaunchKernelGGL(k_bias_silu, dim3(nb((long long)n * kH)), dim3(256), 0, s, m->d_Z1, P + oB1, n, kH, m->d_H1);
    gemm(s, n, kH, kH, 1.f, m->d_H1, kH, 1, P + oW2, kH, 1, 0.f, m->d_Z2, kH, nullptr);
    hipLaunchKernelGGL(k_bias_silu, dim3(nb((long long)n * kH)), dim3(256), 0, s, m->d_Z2, P + oB2, n, kH, m->d_H2);
    gemm(s, n, kOut, kH, 1.f, m->d_H2, kH, 1, P + oW3, kH, 1, 0.f, m->d_Y, kOut, nullptr);
    hipLaunchKernelGGL(k_bias, dim3(nb((long long)n * kOut)), dim3(256), 0, s, m->d_Y, P + oB3, n, kOut);
}

// Backward of the mean squared error of scale * Y against d_T into d_g.
void backward(xpbd_mlp* m, int n)
{
    hipStream_t s = m->stream;
    const float* P = m->d_p;
    float* G = m->d_g;
    float* W = m->d_split;
    hipLaunchKernelGGL(k_mse_grad, dim3(nb((long long)n * kOut)), dim3(256), 0, s, m->d_Y, m->d_T, n,
                       1.0f / m->scale, m->d_dY, m->d_lpart);
    // layer 3: dW3[o][i] = sum_r dY[r][o] H2[r][i]; db3; dH2 = dY W3
    gemm(s, kOut, kH, n, 1.f, m->d_dY, 1, kOut, m->d_H2, 1, kH, 0.f, G + oW3, kH, W);
    hipLaunchKernelGGL(k_colsum, dim3(kOut), dim3(256), 0, s, m->d_dY, n, kOut, G + oB3);
    gemm(s, n, kH, kOut, 1.f, m->d_dY, kOut, 1, P + oW3, 1, kH, 0.f, m->d_dH, kH, nullptr);
    hipLaunchKernelGGL(k_dsilu, dim3(nb((long long)n * kH)), dim3(256), 0, s, m->d_dH, m->d_Z2, n * kH);
    // layer 2
    gemm(s, kH, kH, n, 1.f, m->d_dH, 1, kH, m->d_H1, 1, kH, 0.f, G + oW2, kH, W);
    hipLaunchKernelGGL(k_colsum, dim3(kH), dim3(256), 0, s, m->d_dH, n, kH, G + oB2);
    gemm(s, n, kH, kH, 1.f, m->d_dH, kH, 1, P + oW2, 1, kH, 0.f, m->d_dH1, kH, nullptr);
    hipLaunchKernelGGL(k_dsilu, dim3(nb((long long)n * kH)), dim3(256), 0, s, m->d_dH1, m->d_Z1, n * kH);
    // layer 1: dW1[o][k] = sum_r dZ1[r][o] Xs[r][k] (k < 24)
    gemm(s, kH, kIn, n, 1.f, m->d_dH1, 1, kH, m->d_Xs, 1, kInP, 0.f, G + oW1, kIn, W);
    hipLaunchKernelGGL(k_colsum, dim3(kH), dim3(256), 0, s, m->d_dH1, n, kH, G + oB1);
}

double read_loss(xpbd_mlp* m, int n)
{
    std::vector<double> lp(nb((long long)n * kOut));
    (void)hipMemcpyAsync(lp.data(), m->d_lpart, sizeof(double) * lp.size(), hipMemcpyDeviceToHost, m->stream);
    (void)hipStreamSynchronize(m->stream);
    double s = 0;
    for (double v : lp) s += v;
    return s / ((double)n * kOut);
}

}  // namespace

extern "C" int xpbd_mlp_num_params(int32_t* out)
{
    if (!out) return XPBD_E_INVALID_ARG;
    *out = kNParams;
    return XPBD_OK;
}

extern "C" int xpbd_mlp_create(xpbd_mlp** out)
{
    if (!out) return XPBD_E_INVALID_ARG;
    *out = nullptr;
    xpbd_mlp* m = new (std::nothrow) xpbd_mlp();
    if (!m) return XPBD_E_NOMEM;
    m->version = next_mlp_version();
    hipError_t e = hipGetDevice(&m->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    auto A = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    A(reinterpret_cast<void**>(&m->d_p), sizeof(float) * kNParams);
    A(reinterpret_cast<void**>(&m->d_g), sizeof(float) * kNParams);
    A(reinterpret_cast<void**>(&m->d_m), sizeof(float) * kNParams);
    A(reinterpret_cast<void**>(&m->d_v), sizeof(float) * kNParams);
    A(reinterpret_cast<void**>(&m->d_w1), sizeof(__bf16) * kH * kInP);
    A(reinterpret_cast<void**>(&m->d_w2), sizeof(__bf16) * kH * kH);
    A(reinterpret_cast<void**>(&m->d_w3), sizeof(__bf16) * kOutP * kH);
    A(reinterpret_cast<void**>(&m->d_mean), sizeof(float) * kIn);
    A(reinterpret_cast<void**>(&m->d_istd), sizeof(float) * kIn);
    A(reinterpret_cast<void**>(&m->d_split), sizeof(float) * kMaxSplit * kH * kH);
    if (e != hipSuccess) {
        xpbd_mlp_destroy(m);
        return XPBD_E_NOMEM;
    }
    // Glorot-uniform initial weights, zero biases (fixed seed), identity scaling
    std::vector<float> p(kNParams, 0.f);
    std::mt19937 rng(20099u);
    auto init = [&](int off, int fan_out, int fan_in) {
        const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
        std::uniform_real_distribution<float> U(-a, a);
        for (int i = 0; i < fan_out * fan_in; i++) p[off + i] = U(rng);
    };
    init(oW1, kH, kIn);
    init(oW2, kH, kH);
    init(oW3, kOut, kH);
    std::vector<float> zero(kNParams, 0.f), mean(kIn, 0.f), one(kIn, 1.f);
    e = hipMemcpy(m->d_p, p.data(), sizeof(float) * kNParams, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->d_m, zero.data(), sizeof(float) * kNParams, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->d_v, zero.data(), sizeof(float) * kNParams, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->d_g, zero.data(), sizeof(float) * kNParams, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->d_mean, mean.data(), sizeof(float) * kIn, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->d_istd, one.data(), sizeof(float) * kIn, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = pack(m);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) {
        xpbd_mlp_destroy(m);
        return XPBD_E_HIP;
    }
    *out = m;
    return XPBD_OK;
}

extern "C" int xpbd_mlp_destroy(xpbd_mlp* m)
{
    if (!m) return XPBD_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    free_ws(m);
    void* bufs[] = { m->d_p, m->d_g, m->d_m, m->d_v, m->d_w1, m->d_w2, m->d_w3, m->d_mean, m->d_istd, m->d_split };
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return XPBD_OK;
}

extern "C" const char* xpbd_mlp_last_error(const xpbd_mlp* m) { return m ? m->err.c_str() : "null network"; }

extern "C" int xpbd_mlp_set_params(xpbd_mlp* m, const float* params)
{
    if (!m || !params) return XPBD_E_INVALID_ARG;
    for (int i = 0; i < kNParams; i++)
        if (!std::isfinite(params[i])) return merr(m, XPBD_E_NONFINITE, "non-finite parameter");
    hipError_t e = hipMemcpyAsync(m->d_p, params, sizeof(float) * kNParams, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = pack(m);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    return e == hipSuccess ? XPBD_OK : merr(m, XPBD_E_HIP, hipGetErrorString(e));
}

extern "C" int xpbd_mlp_get_params(xpbd_mlp* m, float* params)
{
    if (!m || !params) return XPBD_E_INVALID_ARG;
    hipError_t e = hipMemcpyAsync(params, m->d_p, sizeof(float) * kNParams, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    return e == hipSuccess ? XPBD_OK : merr(m, XPBD_E_HIP, hipGetErrorString(e));
}

extern "C" int xpbd_mlp_set_scaling(xpbd_mlp* m, const float* mean_in, const float* std_in, float scale_out)
{
    if (!m || !mean_in || !std_in || !(scale_out > 0) || !std::isfinite(scale_out)) return XPBD_E_INVALID_ARG;
    float istd[kIn];
    for (int k = 0; k < kIn; k++) {
        if (!(std_in[k] > 0) || !std::isfinite(std_in[k]) || !std::isfinite(mean_in[k]))
            return merr(m, XPBD_E_INVALID_ARG, "std must be positive and finite");
        istd[k] = 1.0f / std_in[k];
    }
    hipError_t e = hipMemcpyAsync(m->d_mean, mean_in, sizeof(float) * kIn, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(m->d_istd, istd, sizeof(float) * kIn, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    m->scale = scale_out;
    m->version = next_mlp_version();
    if (e == hipSuccess) {
        std::memcpy(m->h_mean, mean_in, sizeof(float) * kIn);
        std::memcpy(m->h_std, std_in, sizeof(float) * kIn);
        m->has_scaling = true;
    }
    return e == hipSuccess ? XPBD_OK : merr(m, XPBD_E_HIP, hipGetErrorString(e));
}

namespace {
constexpr char kMlpMagic[8] = { 'X', 'P', 'B', 'D', 'M', 'L', 'P', '1' };
}

extern "C" int xpbd_mlp_save(xpbd_mlp* m, const char* path)
{
    if (!m || !path) return XPBD_E_INVALID_ARG;
    if (!m->has_scaling) return merr(m, XPBD_E_INVALID_ARG, "no feature scaling set");
    std::vector<float> p(kNParams);
    int rc = xpbd_mlp_get_params(m, p.data());
    if (rc) return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f) return merr(m, XPBD_E_INVALID_ARG, std::string("cannot write ") + path);
    const int32_t sizes[2] = { kNParams, kIn };
    bool ok = std::fwrite(kMlpMagic, 1, sizeof(kMlpMagic), f) == sizeof(kMlpMagic) &&
              std::fwrite(sizes, sizeof(int32_t), 2, f) == 2 &&
              std::fwrite(p.data(), sizeof(float), kNParams, f) == (size_t)kNParams &&
              std::fwrite(m->h_mean, sizeof(float), kIn, f) == (size_t)kIn &&
              std::fwrite(m->h_std, sizeof(float), kIn, f) == (size_t)kIn &&
              std::fwrite(&m->scale, sizeof(float), 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? XPBD_OK : merr(m, XPBD_E_INVALID_ARG, std::string("short write to ") + path);
}

extern "C" int xpbd_mlp_load(xpbd_mlp* m, const char* path)
{
    if (!m || !path) return XPBD_E_INVALID_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return merr(m, XPBD_E_INVALID_ARG, std::string("cannot read ") + path);
    char magic[8];
    int32_t sizes[2];
    std::vector<float> p(kNParams);
    float mean[kIn], sd[kIn], scale = 0.f;
    bool ok = std::fread(magic, 1, sizeof(magic), f) == sizeof(magic) &&
              std::memcmp(magic, kMlpMagic, sizeof(magic)) == 0 && std::fread(sizes, sizeof(int32_t), 2, f) == 2 &&
              sizes[0] == kNParams && sizes[1] == kIn &&
              std::fread(p.data(), sizeof(float), kNParams, f) == (size_t)kNParams &&
              std::fread(mean, sizeof(float), kIn, f) == (size_t)kIn &&
              std::fread(sd, sizeof(float), kIn, f) == (size_t)kIn && std::fread(&scale, sizeof(float), 1, f) == 1;
    char extra;
    ok = ok && std::fread(&extra, 1, 1, f) == 0;  // nothing after the scale
    std::fclose(f);
    if (!ok) return merr(m, XPBD_E_INVALID_ARG, std::string("not a network file of this layout: ") + path);
    int rc = xpbd_mlp_set_params(m, p.data());
    return rc ? rc : xpbd_mlp_set_scaling(m, mean, sd, scale);
}

namespace {
// fused inference on device rows: out[r] = scale * MLP(standardised X[r])
hipError_t mlp_infer(xpbd_mlp* m, hipStream_t s, const float* dX, int n, float4* dout)
{
    if (n <= 0) return hipSuccess;
    const int tiles = (n + kFwdRows - 1) / kFwdRows;
    const int grid = tiles < 512 ? tiles : 512;
    hipLaunchKernelGGL(k_mlp_fused, dim3(grid), dim3(256), 0, s, dX, n, m->d_w1, m->d_w2, m->d_w3, m->d_p, m->d_mean,
                       m->d_istd, m->scale, dout, nullptr);
    return hipGetLastError();
}
}  // namespace

extern "C" int xpbd_mlp_forward(xpbd_mlp* m, int32_t n, const float* x, float* out)
{
    if (!m || n < 0 || (n > 0 && (!x || !out))) return XPBD_E_INVALID_ARG;
    if (n == 0) return XPBD_OK;
    hipError_t e = ensure_ws(m, n);
    if (e != hipSuccess) return merr(m, XPBD_E_NOMEM, hipGetErrorString(e));
    e = hipMemcpyAsync(m->d_X, x, sizeof(float) * kIn * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = mlp_infer(m, m->stream, m->d_X, n, m->d_out);
    std::vector<float4> o(n);
    if (e == hipSuccess)
        e = hipMemcpyAsync(o.data(), m->d_out, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) return merr(m, XPBD_E_HIP, hipGetErrorString(e));
    for (int i = 0; i < n; i++) {
        out[3 * (size_t)i] = o[i].x;
        out[3 * (size_t)i + 1] = o[i].y;
        out[3 * (size_t)i + 2] = o[i].z;
    }
    return XPBD_OK;
}

// ---------------------------------------------------------------------------
// Forward without allocation: the rows are staged through a fixed file-scope
// device array instead of the network's workspace (which xpbd_mlp_forward grows
// with hipMalloc).  Each GPU holds its own copy of the array; its address on a
// device is looked up once (hipGetSymbolAddress with that device current) and
// kept, since the array's name in host code is not a device address.
// ---------------------------------------------------------------------------
namespace {
constexpr int kNoAllocRows = XPBD_MLP_NOALLOC_MAX_ROWS;
// per row: 24 input floats, a float4 network output, 3 packed output floats
constexpr size_t kScratchFloats = (size_t)kNoAllocRows * (kIn + 4 + kOut);
static __device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

// The calling thread's device is made `device` for the lookup and restored.
hipError_t scratch_on(int device, float** out)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) e = hipSetDevice(device);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// out3[3 i + k] = out4[i].k for the first three components
__global__ void k_pack3(const float4* __restrict__ out4, int n, float* __restrict__ out3)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 v = out4[i];
    out3[3 * (size_t)i] = v.x;
    out3[3 * (size_t)i + 1] = v.y;
    out3[3 * (size_t)i + 2] = v.z;
}
}  // namespace

extern "C" int xpbd_mlp_forward_noalloc(xpbd_mlp* m, int32_t n, const float* x, float* out)
{
    if (!m || n < 0 || (n > 0 && (!x || !out))) return XPBD_E_INVALID_ARG;
    if (n > kNoAllocRows)
        return merr(m, XPBD_E_INVALID_ARG,
                    "xpbd_mlp_forward_noalloc: " + std::to_string(n) + " rows exceed the fixed scratch of " +
                        std::to_string(kNoAllocRows) + " rows (call it in chunks, or use xpbd_mlp_forward)");
    if (n == 0) return XPBD_OK;
    float* base = nullptr;
    hipError_t e = scratch_on(m->device, &base);
    if (e != hipSuccess) return merr(m, XPBD_E_HIP, std::string("scratch lookup: ") + hipGetErrorString(e));
    float* dX = base;                                                                   // n x 24
    float4* d4 = reinterpret_cast<float4*>(base + (size_t)kNoAllocRows * kIn);         // n float4
    float* d3 = base + (size_t)kNoAllocRows * (kIn + 4);                                // n x 3
    e = hipMemcpyAsync(dX, x, sizeof(float) * kIn * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = mlp_infer(m, m->stream, dX, n, d4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_pack3, dim3(nb(n)), dim3(256), 0, m->stream, d4, n, d3);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d3, sizeof(float) * kOut * (size_t)n, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    return e == hipSuccess ? XPBD_OK : merr(m, XPBD_E_HIP, hipGetErrorString(e));
}

extern "C" int xpbd_mlp_loss_grad(xpbd_mlp* m, int32_t n, const float* x, const float* y, double* loss, float* grad)
{
    if (!m || n <= 0 || !x || !y || !loss) return XPBD_E_INVALID_ARG;
    hipError_t e = ensure_ws(m, n);
    if (e != hipSuccess) return merr(m, XPBD_E_NOMEM, hipGetErrorString(e));
    e = hipMemcpyAsync(m->d_X, x, sizeof(float) * kIn * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(m->d_T, y, sizeof(float) * kOut * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e != hipSuccess) return merr(m, XPBD_E_HIP, hipGetErrorString(e));
    forward_train(m, n);
    backward(m, n);
    e = hipGetLastError();
    if (e != hipSuccess) return merr(m, XPBD_E_HIP, hipGetErrorString(e));
    *loss = read_loss(m, n);
    if (grad) {
        e = hipMemcpyAsync(grad, m->d_g, sizeof(float) * kNParams, hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    }
    return e == hipSuccess ? XPBD_OK : merr(m, XPBD_E_HIP, hipGetErrorString(e));
}

extern "C" int xpbd_mlp_train(xpbd_mlp* m, int32_t n, const float* x, const float* y, int32_t batch, int32_t epochs,
                              float lr, uint64_t seed, double* final_loss)
{
    if (!m || n <= 0 || !x || !y || batch <= 0 || epochs < 0 || !(lr > 0) || !std::isfinite(lr))
        return XPBD_E_INVALID_ARG;
    if (batch > n) batch = n;
    // all samples live on the device; each epoch gathers minibatches through a
    // seeded permutation
    float *dXall = nullptr, *dYall = nullptr;
    int* dperm = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&dXall), sizeof(float) * kIn * (size_t)n);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&dYall), sizeof(float) * kOut * (size_t)n);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&dperm), sizeof(int) * (size_t)n);
    if (e == hipSuccess) e = ensure_ws(m, batch);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dXall, x, sizeof(float) * kIn * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dYall, y, sizeof(float) * kOut * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(m->stream);
        if (dXall) (void)hipFree(dXall);
        if (dYall) (void)hipFree(dYall);
        if (dperm) (void)hipFree(dperm);
        return merr(m, XPBD_E_NOMEM, hipGetErrorString(e));
    }
    std::mt19937_64 rng(seed);
    std::vector<int32_t> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double last = 0;
    const int nbatch = n / batch > 0 ? n / batch : 1;
    for (int ep = 0; ep < epochs && e == hipSuccess; ep++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        e = hipMemcpyAsync(dperm, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, m->stream);
        for (int bi = 0; bi < nbatch && e == hipSuccess; bi++) {
            hipLaunchKernelGGL(k_gather, dim3(nb((long long)batch * kIn)), dim3(256), 0, m->stream, dXall, dYall,
                               dperm, bi * batch, n, batch, m->d_X, m->d_T);
            forward_train(m, batch);
            backward(m, batch);
            m->adam_t++;
            const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
            hipLaunchKernelGGL(k_adam, dim3(nb(kNParams)), dim3(256), 0, m->stream, m->d_p, m->d_g, m->d_m, m->d_v,
                               kNParams, lr, b1, b2, c1, c2, eps);
            e = hipGetLastError();
        }
        // the host permutation is rewritten next epoch: wait for its copy
        if (e == hipSuccess) last = read_loss(m, batch);
    }
    if (e == hipSuccess) e = pack(m);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipStreamSynchronize(m->stream);
    (void)hipFree(dXall);
    (void)hipFree(dYall);
    (void)hipFree(dperm);
    if (final_loss) *final_loss = last;
    return e == hipSuccess ? XPBD_OK : merr(m, XPBD_E_HIP, hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// learned step on a solver handle
// ---------------------------------------------------------------------------
namespace xpbd {

struct Learned {
    xpbd_mlp* mlp = nullptr;
    int nv = 0;
    int* d_nb_off = nullptr;
    int* d_nb = nullptr;
    float2* d_vmat = nullptr;
    float* d_pinfrac = nullptr;
    float* d_feat = nullptr;     // nv x 24
    float4* d_pred = nullptr;    // in-step motion after the prediction
    float4* d_cprev = nullptr;   // previous step's solver correction
    float4* d_corr = nullptr;    // network output
    float4* d_save = nullptr;    // snapshot for the fallback: pos (2 nv), lo (nv), vel (nv)
    unsigned long long* d_step_save = nullptr;
    double* d_res = nullptr;     // squared residual of the learned solve
    // Captured learned steps, one per (network version, handle graph
    // generation, detection, iterations, start mode).
    struct Graph {
        uint64_t net = 0, gen = 0;
        int det = 0, iters = 0, start = 0;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;
    // inference timing (kernel timing mode of the handle)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double mlp_ms = 0;
    int64_t mlp_calls = 0;
    int64_t steps = 0, fallbacks = 0;
    // XPBD_LEARNED_START_*: how the solve after the network's correction starts
    int start = XPBD_LEARNED_START_MULTIPLIERS;
};

static void free_learned(Learned* L)
{
    if (!L) return;
    for (auto& g : L->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    void* bufs[] = { L->d_nb_off, L->d_nb, L->d_vmat, L->d_pinfrac, L->d_feat, L->d_pred,
                     L->d_cprev,  L->d_corr, L->d_save, L->d_step_save, L->d_res };
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (L->ev0) (void)hipEventDestroy(L->ev0);
    if (L->ev1) (void)hipEventDestroy(L->ev1);
    delete L;
}

void destroy_learned(xpbd_handle* h)
{
    if (!h || !h->learned) return;
    (void)hipStreamSynchronize(h->stream);
    free_learned(h->learned);
    h->learned = nullptr;
}

int learned_reset_history(xpbd_handle* h)
{
    if (!h->learned) return XPBD_OK;
    hipError_t e = launch_fill_zero(h->stream, h->learned->d_cprev, sizeof(float4) * (size_t)h->nv);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "learned history reset");
}

// Device buffers of the learned step (1-ring CSR, material summary, scratch).
static int ensure_learned(xpbd_handle* h)
{
    if (h->learned) return XPBD_OK;
    const int nv = h->nv, nt = h->nt;
    // 1-ring neighbours from the element list (sorted, unique)
    std::vector<int64_t> pairs;
    pairs.reserve(12 * (size_t)nt);
    for (int e = 0; e < nt; e++) {
        const int32_t* t = &h->tets_host[4 * (size_t)e];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++)
                if (a != b) pairs.push_back(((int64_t)t[a] << 32) | (uint32_t)t[b]);
    }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    std::vector<int32_t> off(nv + 1, 0), nbr(pairs.size());
    for (size_t k = 0; k < pairs.size(); k++) {
        off[(pairs[k] >> 32) + 1]++;
        nbr[k] = (int32_t)(pairs[k] & 0xffffffff);
    }
    for (int v = 0; v < nv; v++) off[v + 1] += off[v];
    std::vector<float> pinfrac(nv, 0.f);
    for (int v = 0; v < nv; v++) {
        int c = 0;
        for (int k = off[v]; k < off[v + 1]; k++) c += h->pinned[nbr[k]] ? 1 : 0;
        pinfrac[v] = off[v + 1] > off[v] ? (float)c / (float)(off[v + 1] - off[v]) : 0.f;
    }
    Learned* L = new (std::nothrow) Learned();
    if (!L) return set_err(h, XPBD_E_NOMEM, "learned state allocation failed");
    L->nv = nv;
    hipError_t e = hipSuccess;
    auto A = [&](auto** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), bytes > 0 ? bytes : 16);
    };
    A(&L->d_nb_off, sizeof(int) * (nv + 1));
    A(&L->d_nb, sizeof(int) * nbr.size());
    A(&L->d_vmat, sizeof(float2) * nv);
    A(&L->d_pinfrac, sizeof(float) * nv);
    A(&L->d_feat, sizeof(float) * kIn * (size_t)nv);
    A(&L->d_pred, sizeof(float4) * nv);
    A(&L->d_cprev, sizeof(float4) * nv);
    A(&L->d_corr, sizeof(float4) * nv);
    A(&L->d_save, sizeof(float4) * 4 * (size_t)nv);
    A(&L->d_step_save, sizeof(unsigned long long));
    A(&L->d_res, sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(L->d_nb_off, off.data(), sizeof(int) * (nv + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && !nbr.empty())
        e = hipMemcpy(L->d_nb, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->d_vmat, h->vmat.data(), sizeof(float2) * nv, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->d_pinfrac, pinfrac.data(), sizeof(float) * nv, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&L->ev0);
    if (e == hipSuccess) e = hipEventCreate(&L->ev1);
    if (e == hipSuccess) e = launch_fill_zero(h->stream, L->d_cprev, sizeof(float4) * nv);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        free_learned(L);
        return hip_err(h, e, "learned set-up");
    }
    h->learned = L;
    return XPBD_OK;
}

static int enqueue_features(xpbd_handle* h)
{
    Learned* L = h->learned;
    hipLaunchKernelGGL(k_features, dim3(nb(h->nv)), dim3(256), 0, h->stream, h->d_pos, h->d_vel, L->d_cprev,
                       L->d_nb_off, L->d_nb, L->d_vmat, L->d_pinfrac, h->nv, L->d_feat);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "features launch");
}

static int enqueue_copy_pred(xpbd_handle* h)
{
    hipLaunchKernelGGL(k_copy_delta, dim3(nb(h->nv)), dim3(256), 0, h->stream, h->d_pos, h->learned->d_pred, h->nv);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "copy launch");
}

static int enqueue_store_cprev(xpbd_handle* h)
{
    Learned* L = h->learned;
    hipLaunchKernelGGL(k_store_cprev, dim3(nb(h->nv)), dim3(256), 0, h->stream, h->d_pos, h->d_vel, L->d_pred,
                       L->d_cprev, h->nv);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "correction launch");
}

// Multiplier start of the learned step, on the corrected positions: the
// contacts are projected once from zero multipliers (a Gauss-Seidel contact
// pass, or the Jacobi contact sum), so that the start does not penetrate the
// collider, then every element's multipliers are set to -C / alpha~ there.
// Attachment multipliers start at zero.  The iterations that follow keep all
// of them.
static int enqueue_multiplier_start(xpbd_handle* h)
{
    hipStream_t s = h->stream;
    hipError_t e = hipSuccess;
    if (h->natt > 0) e = launch_fill_zero(s, h->d_att_lam, sizeof(double) * (size_t)h->natt);
    if (e == hipSuccess && contacts_enabled(h)) {
        if (h->params.solver == XPBD_SOLVER_GS_COLORED) {
            e = launch_contacts_gs(s, h->d_faces, h->d_fci_off, h->nfcolors, h->ninst, h->d_cpos, h->d_clist,
                                   h->d_rest, h->d_pos, h->d_lo, h->d_cp, h->d_cn, h->d_cb, h->d_clam, true);
        } else {
            int rc = enqueue_contacts_jacobi(h, true);
            if (rc) return rc;
            e = launch_jacobi_apply(s, h->d_pos, h->d_acc, h->nv);
        }
    }
    if (e == hipSuccess) e = launch_lambda_start(s, tet_arrays(h), h->d_pos, h->d_lo, h->nt);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "multiplier start");
}

// Consistent start (XPBD_LEARNED_START_CONSISTENT, the oracle's orc_step_learned):
// the multipliers lambda0 = -C / alpha~ at the corrected positions xl = xt + corr,
// then the positions x0 = xt + M^-1 sum grad C(xl)^T lambda0 from the kept
// prediction, so that the momentum balance M (x - xt) = grad C^T lambda, which
// the XPBD updates preserve, holds at the first iteration.  Contact and
// attachment multipliers start at zero, as in a plain solve.
static int enqueue_consistent_start(xpbd_handle* h)
{
    hipStream_t s = h->stream;
    hipError_t e = hipSuccess;
    if (h->natt > 0) e = launch_fill_zero(s, h->d_att_lam, sizeof(double) * (size_t)h->natt);
    if (e == hipSuccess && h->nf > 0) e = launch_fill_zero(s, h->d_clam, sizeof(double) * (size_t)h->nf);
    const TetArrays ta = tet_arrays(h);
    if (e == hipSuccess) e = launch_lambda_start(s, ta, h->d_pos, h->d_lo, h->nt);
    if (e == hipSuccess) e = launch_element_forces(s, ta, element_force_args(h));
    if (e == hipSuccess) e = launch_consistent_shift(s, h->d_pos, h->learned->d_pred, h->d_vt_off, h->d_eforce, h->nv);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "consistent start");
}

}  // namespace xpbd

using namespace xpbd;

extern "C" int xpbd_attach_mlp(xpbd_handle* h, xpbd_mlp* m)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (!m) {
        if (h->learned) h->learned->mlp = nullptr;
        return XPBD_OK;
    }
    int rc = ensure_learned(h);
    if (rc) return rc;
    h->learned->mlp = m;
    return XPBD_OK;
}

extern "C" int xpbd_features(xpbd_handle* h, float* out)
{
    if (!h || !out) return h ? set_err(h, XPBD_E_INVALID_ARG, "out is NULL") : XPBD_E_INVALID_ARG;
    int rc = ensure_learned(h);
    if (rc) return rc;
    // features of the state after this step's prediction; the prediction is undone afterwards
    if ((rc = enqueue_predict(h)) || (rc = enqueue_features(h))) return rc;
    hipLaunchKernelGGL(k_zero_delta, dim3(nb(h->nv)), dim3(256), 0, h->stream, h->d_pos, h->nv);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, h->learned->d_feat, sizeof(float) * kIn * (size_t)h->nv, hipMemcpyDeviceToHost,
                           h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "features");
}

extern "C" int xpbd_collect_training(xpbd_handle* h, int32_t nsteps, int32_t cap, float* x, float* y, int32_t* rows)
{
    if (!h || nsteps < 0 || cap < 0 || !rows || (cap > 0 && (!x || !y)))
        return h ? set_err(h, XPBD_E_INVALID_ARG, "collect_training: bad arguments") : XPBD_E_INVALID_ARG;
    int rc = ensure_learned(h);
    if (rc) return rc;
    Learned* L = h->learned;
    const int nv = h->nv;
    std::vector<float> feat((size_t)kIn * nv);
    std::vector<float4> corr(nv);
    int32_t count = 0;
    for (int32_t s = 0; s < nsteps; s++) {
        const bool det = detect_this_step(h);
        if (det && (rc = enqueue_detect(h))) return rc;
        if ((rc = enqueue_predict(h)) || (rc = enqueue_copy_pred(h)) || (rc = enqueue_features(h))) return rc;
        hipError_t e = hipMemcpyAsync(feat.data(), L->d_feat, sizeof(float) * feat.size(), hipMemcpyDeviceToHost,
                                      h->stream);
        if (e != hipSuccess) return hip_err(h, e, "collect features");
        if ((rc = enqueue_iterations(h, h->params.iters, false))) return rc;
        if (records_residuals(h) && (rc = enqueue_step_residual(h))) return rc;
        if ((rc = enqueue_local_collision(h))) return rc;
        if ((rc = enqueue_store_cprev(h)) || (rc = enqueue_finish(h))) return rc;
        e = hipMemcpyAsync(corr.data(), L->d_cprev, sizeof(float4) * nv, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_err(h, e, "collect targets");
        h->host_step++;
        for (int i = 0; i < nv && count < cap; i++) {
            if (h->pinned[i]) continue;
            std::memcpy(x + (size_t)count * kIn, &feat[(size_t)i * kIn], sizeof(float) * kIn);
            y[(size_t)count * kOut] = corr[i].x;
            y[(size_t)count * kOut + 1] = corr[i].y;
            y[(size_t)count * kOut + 2] = corr[i].z;
            count++;
        }
    }
    *rows = count;
    return XPBD_OK;
}

namespace xpbd {
// The learned step up to its end, before the fallback decision: detection (on
// the plain solver's cadence), a snapshot of the start-of-step state, the
// prediction, features, network correction, the multiplier start, iters
// iterations, the squared residual into L->d_res, then the rest of the step
// (EVERY_SUBSTEP residual, local collision passes, the correction history and
// the velocity update).  A fallback restores the snapshot and redoes the
// whole step as the plain solver does, so finishing the step before the
// decision changes nothing on either branch.
static int enqueue_learned_attempt(xpbd_handle* h, int32_t iters, bool det, bool timed)
{
    Learned* L = h->learned;
    const int nv = h->nv;
    hipStream_t s = h->stream;
    int rc;
    if (det && (rc = enqueue_detect(h))) return rc;
    hipError_t e = hipMemcpyAsync(L->d_save, h->d_pos, sizeof(float4) * 2 * (size_t)nv, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(L->d_save + 2 * (size_t)nv, h->d_lo, sizeof(float4) * nv, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(L->d_save + 3 * (size_t)nv, h->d_vel, sizeof(float4) * nv, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(L->d_step_save, h->d_step, sizeof(unsigned long long), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_err(h, e, "learned snapshot");
    if ((rc = enqueue_predict(h)) || (rc = enqueue_copy_pred(h)) || (rc = enqueue_features(h))) return rc;
    if (timed) (void)hipEventRecord(L->ev0, s);
    e = mlp_infer(L->mlp, s, L->d_feat, nv, L->d_corr);
    if (timed) (void)hipEventRecord(L->ev1, s);
    if (e != hipSuccess) return hip_err(h, e, "mlp launch");
    hipLaunchKernelGGL(k_apply_corr, dim3(nb(nv)), dim3(256), 0, s, h->d_pos, h->d_vel, L->d_corr, nv);
    if ((e = hipGetLastError()) != hipSuccess) return hip_err(h, e, "apply launch");
    const bool warm = L->start == XPBD_LEARNED_START_MULTIPLIERS;
    const bool consistent = L->start == XPBD_LEARNED_START_CONSISTENT;
    if (warm && (rc = enqueue_multiplier_start(h))) return rc;
    if (consistent && (rc = enqueue_consistent_start(h))) return rc;
    if ((rc = enqueue_iterations(h, iters, false, !(warm || consistent)))) return rc;
    e = launch_residual(s, tet_arrays(h), h->d_pos, h->d_lo, true, h->nt, h->d_part, h->red_blocks, nullptr);
    if (e == hipSuccess) e = launch_sum_partials(s, h->d_part, h->red_blocks, L->d_res);
    if (e != hipSuccess) return hip_err(h, e, "learned residual");
    if (records_residuals(h) && (rc = enqueue_step_residual(h))) return rc;
    if ((rc = enqueue_local_collision(h))) return rc;
    if ((rc = enqueue_store_cprev(h)) || (rc = enqueue_finish(h))) return rc;
    return XPBD_OK;
}

// Launch the attempt from a captured graph (built on first use of its shape).
static int launch_learned_attempt(xpbd_handle* h, int32_t iters, bool det)
{
    Learned* L = h->learned;
    const uint64_t net = L->mlp->version;
    hipGraphExec_t exec = nullptr;
    for (size_t i = 0; i < L->graphs.size();) {
        Learned::Graph& g = L->graphs[i];
        if (g.gen != h->graph_gen) {  // the handle's captured steps went stale: so did this one
            (void)hipGraphExecDestroy(g.exec);
            L->graphs[i] = L->graphs.back();
            L->graphs.pop_back();
            continue;
        }
        if (g.net == net && g.det == (int)det && g.iters == iters && g.start == L->start) exec = g.exec;
        i++;
    }
    if (!exec) {
        hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_err(h, e, "hipStreamBeginCapture");
        int rc = enqueue_learned_attempt(h, iters, det, false);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(h->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_err(h, e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_err(h, e, "hipGraphInstantiate");
        // graphs of other network versions are unlikely to be replayed again
        for (size_t i = 0; i < L->graphs.size();) {
            if (L->graphs[i].net != net) {
                (void)hipGraphExecDestroy(L->graphs[i].exec);
                L->graphs[i] = L->graphs.back();
                L->graphs.pop_back();
            } else {
                i++;
            }
        }
        L->graphs.push_back({ net, h->graph_gen, (int)det, iters, L->start, exec });
    }
    hipError_t e = hipGraphLaunch(exec, h->stream);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "hipGraphLaunch");
}
}  // namespace xpbd

extern "C" int xpbd_step_learned(xpbd_handle* h, int32_t iters, double ref_residual, double fallback_ratio,
                                 int32_t* fell_back, double* residual)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (iters < 0 || !std::isfinite(ref_residual) || !std::isfinite(fallback_ratio))
        return set_err(h, XPBD_E_INVALID_ARG, "step_learned: iters >= 0 and finite thresholds required");
    if (!h->learned || !h->learned->mlp) return set_err(h, XPBD_E_STATE, "no network attached");
    Learned* L = h->learned;
    const int nv = h->nv;
    hipStream_t s = h->stream;
    const bool det = detect_this_step(h);
    // kernel timing mode launches eagerly and times the inference with events
    int rc = h->timing ? enqueue_learned_attempt(h, iters, det, true) : launch_learned_attempt(h, iters, det);
    if (rc) return rc;
    double r2 = 0;
    hipError_t e = hipMemcpyAsync(&r2, L->d_res, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(h, e, "learned step");
    double r = std::sqrt(r2);
    if (h->timing) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, L->ev0, L->ev1);
        L->mlp_ms += ms;
        L->mlp_calls++;
    }
    const bool fb = ref_residual > 0 && fallback_ratio > 0 && !(r <= fallback_ratio * ref_residual);
    if (fb) {
        // redo the step from the snapshot as the plain solver does
        e = hipMemcpyAsync(h->d_pos, L->d_save, sizeof(float4) * 2 * (size_t)nv, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h->d_lo, L->d_save + 2 * (size_t)nv, sizeof(float4) * nv, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h->d_vel, L->d_save + 3 * (size_t)nv, sizeof(float4) * nv, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h->d_step, L->d_step_save, sizeof(unsigned long long), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_err(h, e, "learned restore");
        if ((rc = enqueue_predict(h)) || (rc = enqueue_copy_pred(h))) return rc;
        if ((rc = enqueue_iterations(h, h->params.iters, false))) return rc;
        if ((rc = device_residual(h, &r, nullptr))) return rc;
        if (records_residuals(h) && (rc = enqueue_step_residual(h))) return rc;
        if ((rc = enqueue_local_collision(h))) return rc;
        if ((rc = enqueue_store_cprev(h)) || (rc = enqueue_finish(h))) return rc;
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_err(h, e, "learned fallback");
        L->fallbacks++;
    }
    h->host_step++;
    L->steps++;
    if (fell_back) *fell_back = fb ? 1 : 0;
    if (residual) *residual = r;
    return XPBD_OK;
}

extern "C" int xpbd_set_learned_start(xpbd_handle* h, int32_t mode)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (mode != XPBD_LEARNED_START_POSITIONS && mode != XPBD_LEARNED_START_MULTIPLIERS &&
        mode != XPBD_LEARNED_START_CONSISTENT)
        return set_err(h, XPBD_E_INVALID_ARG, "unknown learned start mode");
    int rc = ensure_learned(h);
    if (rc) return rc;
    if (mode == XPBD_LEARNED_START_CONSISTENT && (rc = ensure_vertex_slots(h))) return rc;
    h->learned->start = mode;
    return XPBD_OK;
}

extern "C" int xpbd_learned_stats(xpbd_handle* h, int64_t* steps, int64_t* fallbacks, double* mlp_ms,
                                  int64_t* mlp_calls)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (!h->learned) return set_err(h, XPBD_E_STATE, "no network attached");
    if (steps) *steps = h->learned->steps;
    if (fallbacks) *fallbacks = h->learned->fallbacks;
    if (mlp_ms) *mlp_ms = h->learned->mlp_ms;
    if (mlp_calls) *mlp_calls = h->learned->mlp_calls;
    return XPBD_OK;
}
