// Candidate predictor kernels for gfx950: a small MLP over the normalised
// texels of a block (BF16 operands on the matrix cores, FP32 accumulation),
// with the training kernels (backward, softmax cross-entropy, Adam) and the
// inference-side top-K selection.  The reference has no learned component;
// these kernels serve the learned candidate policy of the encoder.
//
// GEMM: C[M x N] (fp32) = op(A)[M x K] * op(B)[K x N] (+ bias[N]) (+ C),
// A and B in bf16.  A 256-thread workgroup computes a 64 x 64 tile; each of
// its four waves owns a 32 x 32 quarter as 2 x 2 v_mfma_f32_16x16x32_bf16
// tiles.  Both operand tiles sit in LDS with K contiguous (row stride 40
// elements to spread the 16-byte fragment reads over the banks).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/astcg_predictor.h"

namespace {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 64, kBN = 64, kBK = 32, kLdsStride = 40;

__device__ __forceinline__ float bf2f(bf16 v) { return (float)v; }

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const bf16* __restrict__ A, int lda,
                                               const bf16* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                               const float* __restrict__ bias, int accumulate) {
    __shared__ __align__(16) bf16 As[kBM * kLdsStride];
    __shared__ __align__(16) bf16 Bt[kBN * kLdsStride];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int bm = blockIdx.y * kBM, bn = blockIdx.x * kBN;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const bf16 zero = (bf16)0.0f;
    for (int bk = 0; bk < K; bk += kBK) {
        // A tile: (m, k) -> As[m][k]
        if (!TA) {
            const int m = t >> 2, k0 = (t & 3) * 8;
            const int gm = bm + m, gk = bk + k0;
            bf16* dst = As + m * kLdsStride + k0;
#pragma unroll
            for (int j = 0; j < 8; ++j) dst[j] = (gm < M && gk + j < K) ? A[(size_t)gm * lda + gk + j] : zero;
        } else {
            const int k = t >> 3, m0 = (t & 7) * 8;
            const int gk = bk + k, gm = bm + m0;
#pragma unroll
            for (int j = 0; j < 8; ++j) As[(m0 + j) * kLdsStride + k] = (gk < K && gm + j < M) ? A[(size_t)gk * lda + gm + j] : zero;
        }
        // B tile: (k, n) -> Bt[n][k]
        if (!TB) {
            const int k = t >> 3, n0 = (t & 7) * 8;
            const int gk = bk + k, gn = bn + n0;
#pragma unroll
            for (int j = 0; j < 8; ++j) Bt[(n0 + j) * kLdsStride + k] = (gk < K && gn + j < N) ? B[(size_t)gk * ldb + gn + j] : zero;
        } else {
            const int n = t >> 2, k0 = (t & 3) * 8;
            const int gn = bn + n, gk = bk + k0;
            bf16* dst = Bt + n * kLdsStride + k0;
#pragma unroll
            for (int j = 0; j < 8; ++j) dst[j] = (gn < N && gk + j < K) ? B[(size_t)gn * ldb + gk + j] : zero;
        }
        __syncthreads();
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            a[i] = *reinterpret_cast<const bf16x8*>(As + (32 * wm + 16 * i + (lane & 15)) * kLdsStride + 8 * (lane >> 4));
            b[i] = *reinterpret_cast<const bf16x8*>(Bt + (32 * wn + 16 * i + (lane & 15)) * kLdsStride + 8 * (lane >> 4));
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // D fragment: lane holds rows 4 * (lane / 16) + r, column lane % 16
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = bn + 32 * wn + 16 * j + (lane & 15);
            if (col >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = bm + 32 * wm + 16 * i + 4 * (lane >> 4) + r;
                if (row >= M) continue;
                float v = acc[i][j][r];
                if (bias) v += bias[col];
                float* c = C + (size_t)row * ldc + col;
                *c = accumulate ? *c + v : v;
            }
        }
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

__global__ void k_gelu_fwd(const float* z, bf16* h, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) h[i] = (bf16)gelu(z[i]);
}

__global__ void k_gelu_bwd(const float* dh, const float* z, bf16* dz, float* dz32, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = dh[i] * gelu_grad(z[i]);
    dz[i] = (bf16)g;
    if (dz32) dz32[i] = g;
}

__global__ void k_to_bf16(const float* x, bf16* y, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (bf16)x[i];
}

// Column sums of a rows x cols fp32 matrix (bias gradients).
__global__ void k_colsum(const float* x, int rows, int cols, float* out, int accumulate) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) s += x[(size_t)r * cols + c];
    out[c] = accumulate ? out[c] + s : s;
}

// Softmax cross-entropy over the heads of one row (one wave per row).  Heads
// are consecutive logit ranges; a label < 0 masks its head.  Writes
// dlogits = (softmax - onehot) * scale (bf16 and fp32) and the row loss.
__global__ void __launch_bounds__(64) k_softmax_xent(const float* logits, int rows, int ld, int nheads,
                                                      const int* head_off, const int* head_len, const int* labels,
                                                      float scale, bf16* dlogits, float* dlogits32, float* loss) {
    const int r = blockIdx.x, lane = threadIdx.x;
    if (r >= rows) return;
    const float* x = logits + (size_t)r * ld;
    float row_loss = 0.0f;
    for (int h = 0; h < nheads; ++h) {
        const int off = head_off[h], len = head_len[h], lab = labels[(size_t)r * nheads + h];
        float mx = -3.4e38f;
        for (int i = lane; i < len; i += 64) mx = fmaxf(mx, x[off + i]);
        for (int m = 32; m > 0; m >>= 1) mx = fmaxf(mx, __shfl_xor(mx, m));
        float se = 0.0f;
        for (int i = lane; i < len; i += 64) se += expf(x[off + i] - mx);
        for (int m = 32; m > 0; m >>= 1) se += __shfl_xor(se, m);
        const float lse = mx + logf(se);
        for (int i = lane; i < len; i += 64) {
            const size_t o = (size_t)r * ld + off + i;
            float g = 0.0f;
            if (lab >= 0) g = (expf(x[off + i] - lse) - (i == lab ? 1.0f : 0.0f)) * scale;
            dlogits[o] = (bf16)g;
            if (dlogits32) dlogits32[o] = g;
        }
        if (lab >= 0) row_loss += lse - x[off + lab];
    }
    if (lane == 0 && loss) loss[r] = row_loss;
}

// Softmax cross-entropy restricted to candidate columns, one lane per row:
// group g of row r is the columns cols[(r * ngroups + g) * ncand + j]
// (< 0 = absent) and target[r * ngroups + g] the correct j (< 0 = none).
// Adds (softmax - onehot) * scale at those columns into dlogits (bf16) and
// dlogits32 (fp32, when not null), and the groups' negative log-likelihoods
// into loss[r] (when not null).  A column named twice gets both terms.
__global__ void k_rank_xent(const float* logits, int rows, int ld, int ngroups, int ncand, const int* cols,
                            const int* target, float scale, bf16* dlogits, float* dlogits32, float* loss) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const float* x = logits + (size_t)r * ld;
    float row_loss = 0.0f;
    for (int g = 0; g < ngroups; ++g) {
        const int t = target[(size_t)r * ngroups + g];
        const int* c = cols + ((size_t)r * ngroups + g) * ncand;
        if (t < 0 || t >= ncand || c[t] < 0) continue;
        float mx = -3.4e38f;
        for (int j = 0; j < ncand; ++j)
            if (c[j] >= 0) mx = fmaxf(mx, x[c[j]]);
        float se = 0.0f;
        for (int j = 0; j < ncand; ++j)
            if (c[j] >= 0) se += expf(x[c[j]] - mx);
        const float lse = mx + logf(se);
        for (int j = 0; j < ncand; ++j) {
            if (c[j] < 0) continue;
            const float gr = (expf(x[c[j]] - lse) - (j == t ? 1.0f : 0.0f)) * scale;
            const size_t o = (size_t)r * ld + c[j];
            dlogits[o] = (bf16)((float)dlogits[o] + gr);
            if (dlogits32) dlogits32[o] += gr;
        }
        row_loss += lse - x[c[t]];
    }
    if (loss) loss[r] += row_loss;
}

// Adam (bias-corrected) on fp32 master weights; refreshes the bf16 copy.
__global__ void k_adam(float* w, const float* g, float* m, float* v, bf16* w16, long long n, float lr, float b1,
                       float b2, float eps, float c1, float c2, float decay) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + decay * w[i]);
    w[i] = wi;
    if (w16) w16[i] = (bf16)wi;
}

// Block texels -> normalised RGB features (u8 / 255), one row per block;
// edge texels clamp to the image as the encoder does.
__global__ void k_features(const uint8_t* rgba, int width, int height, int bx, int by, int blk0, int nblk,
                           bf16* x, int ldx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int tc = bx * by;
    if (i >= nblk * tc) return;
    const int b = i / tc, t = i - b * tc;
    const int g = blk0 + b;
    const int blocks_x = (width + bx - 1) / bx;
    const int x0 = (g % blocks_x) * bx, y0 = (g / blocks_x) * by;
    const int px = min(x0 + t % bx, width - 1), py = min(y0 + t / bx, height - 1);
    const uint8_t* p = rgba + ((size_t)py * width + px) * 4;
    bf16* row = x + (size_t)b * ldx + 3 * t;
    for (int c = 0; c < 3; ++c) row[c] = (bf16)((float)p[c] * (1.0f / 255.0f));
}

// Top-k indices of each head of each row (one wave per row and head):
// repeated arg-max with the lowest index winning ties.
__global__ void __launch_bounds__(64) k_topk(const float* logits, int rows, int ld, int nheads, const int* head_off,
                                             const int* head_len, int kmax, const int* head_k, int* out_idx,
                                             float* out_prob) {
    const int r = blockIdx.x, h = blockIdx.y, lane = threadIdx.x;
    if (r >= rows || h >= nheads) return;
    const int k = head_k ? min(head_k[h], kmax) : kmax;
    if (k <= 0) return;
    const float* x = logits + (size_t)r * ld + head_off[h];
    const int len = head_len[h];
    float mx = -3.4e38f;
    for (int i = lane; i < len; i += 64) mx = fmaxf(mx, x[i]);
    for (int m = 32; m > 0; m >>= 1) mx = fmaxf(mx, __shfl_xor(mx, m));
    float se = 0.0f;
    for (int i = lane; i < len; i += 64) se += expf(x[i] - mx);
    for (int m = 32; m > 0; m >>= 1) se += __shfl_xor(se, m);
    float prev_v = 3.4e38f;
    int prev_i = -1;
    for (int j = 0; j < k; ++j) {
        // best value below (prev_v, prev_i) in (value desc, index asc) order
        float bv = -3.4e38f;
        int bi = 0x7FFFFFFF;
        for (int i = lane; i < len; i += 64) {
            const float v = x[i];
            const bool after = v < prev_v || (v == prev_v && i > prev_i);
            if (after && (v > bv || (v == bv && i < bi))) { bv = v; bi = i; }
        }
        for (int m = 32; m > 0; m >>= 1) {
            const float ov = __shfl_xor(bv, m);
            const int oi = __shfl_xor(bi, m);
            if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        if (lane == 0) {
            out_idx[((size_t)r * nheads + h) * kmax + j] = bi < len ? bi : -1;
            if (out_prob) out_prob[((size_t)r * nheads + h) * kmax + j] = bi < len ? expf(bv - mx) / se : 0.0f;
        }
        prev_v = bv;
        prev_i = bi;
    }
}

inline unsigned grid1(long long n, int threads) { return (unsigned)((n + threads - 1) / threads); }


// ---- fused inference for the encoder's search plan -----------------------------
// 32 rows (blocks) per 256-thread workgroup, hidden width 256: features ->
// H1 = GELU(X W1 + b1) -> H2 = GELU(H1 W2 + b2) (bf16, in LDS) -> logits of
// the partition-count and trial-mode heads (LDS) -> softmax top-k of both.
// Weights are stored transposed (N x K, K contiguous) so that a lane's B
// fragment of v_mfma_f32_16x16x32_bf16 is one 16-byte load.
constexpr int kFR = 32, kFH = 256, kFS = kFH + 8;  // rows, hidden width, LDS row stride (bf16)

// out[32][256] = GELU(in[32][K] W + b) in bf16; wave w computes hidden
// columns 64w .. 64w + 63 (4 column tiles) for both 16-row tiles.
template <int K>
__device__ __forceinline__ void dense_gelu(const bf16* in, const bf16* __restrict__ Wt, const float* __restrict__ b,
                                           bf16* out, int wave, int lane) {
    f32x4 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int lr = lane & 15, lk = 8 * (lane >> 4);
#pragma unroll 2
    for (int ks = 0; ks < K / 32; ++ks) {
        bf16x8 a[2], w[4];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(in + (16 * i + lr) * kFS + 32 * ks + lk);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            w[j] = *reinterpret_cast<const bf16x8*>(Wt + (size_t)(64 * wave + 16 * j + lr) * K + 32 * ks + lk);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], w[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = 64 * wave + 16 * j + lr;
        const float bias = b[col];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) out[(16 * i + 4 * (lane >> 4) + r) * kFS + col] = (bf16)gelu(acc[i][j][r] + bias);
    }
}

template <int KIN>
__global__ void __launch_bounds__(256) k_infer(int rows, const bf16* __restrict__ X, int ldx, int nin,
                                                const bf16* __restrict__ W1t, const float* __restrict__ b1,
                                                const bf16* __restrict__ W2t, const float* __restrict__ b2,
                                                const bf16* __restrict__ W3t, const float* __restrict__ b3, int n3p,
                                                int nmodes, int kmax, int kpc, int kmode, int* __restrict__ out_idx,
                                                float* __restrict__ out_prob, bf16* __restrict__ h2_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    bf16* sA = reinterpret_cast<bf16*>(smem);              // [32][kFS]: X, then H2
    bf16* sB = sA + kFR * kFS;                             // [32][kFS]: H1
    float* sL = reinterpret_cast<float*>(sB + kFR * kFS);  // [32][n3p + 1]: logits
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r0 = blockIdx.x * kFR;
    for (int i = t; i < kFR * KIN; i += 256) {
        const int r = i / KIN, k = i - r * KIN;
        sA[r * kFS + k] = (r0 + r < rows && k < nin) ? X[(size_t)(r0 + r) * ldx + k] : (bf16)0.0f;
    }
    __syncthreads();
    dense_gelu<KIN>(sA, W1t, b1, sB, wave, lane);
    __syncthreads();
    dense_gelu<kFH>(sB, W2t, b2, sA, wave, lane);
    __syncthreads();
    // output layer: column tiles wave, wave + 4, ... of n3p / 16
    {
        const int nt = n3p / 16, lr = lane & 15, lk = 8 * (lane >> 4), ls = n3p + 1;
        f32x4 acc[2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int ks = 0; ks < kFH / 32; ++ks) {
            bf16x8 a[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(sA + (16 * i + lr) * kFS + 32 * ks + lk);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int tile = wave + 4 * j;
                if (tile >= nt) continue;
                const bf16x8 w = *reinterpret_cast<const bf16x8*>(W3t + (size_t)(16 * tile + lr) * kFH + 32 * ks + lk);
#pragma unroll
                for (int i = 0; i < 2; ++i) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], w, acc[i][j], 0, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int tile = wave + 4 * j;
            if (tile >= nt) continue;
            const int col = 16 * tile + lr;
            const float bias = b3[col];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) sL[(16 * i + 4 * (lane >> 4) + r) * ls + col] = acc[i][j][r] + bias;
        }
    }
    if (h2_out)
        for (int i = t; i < kFR * (kFH / 8); i += 256) {
            const int r = i / (kFH / 8), c = 8 * (i % (kFH / 8));
            if (r0 + r < rows)
                *reinterpret_cast<bf16x8*>(h2_out + (size_t)(r0 + r) * kFH + c) =
                    *reinterpret_cast<const bf16x8*>(sA + r * kFS + c);
        }
    __syncthreads();
    // top-k, 8 lanes per row: rank of each column = columns with a larger
    // logit, or an equal one at a lower index
    const int r = t >> 3, sub = t & 7;
    const float* l = sL + r * (n3p + 1);
    const bool live = r0 + r < rows;
    const size_t o = (size_t)(r0 + r) * 5 * kmax;
    if (live && sub == 0) {
        float mx = l[0];
        for (int j = 1; j < 4; ++j) mx = fmaxf(mx, l[j]);
        float se = 0.0f;
        for (int j = 0; j < 4; ++j) se += expf(l[j] - mx);
        for (int j = 0; j < 4; ++j) {
            int rk = 0;
            for (int i = 0; i < 4; ++i) rk += (l[i] > l[j]) || (l[i] == l[j] && i < j);
            if (rk < kpc) {
                out_idx[o + rk] = j;
                if (out_prob) out_prob[o + rk] = expf(l[j] - mx) / se;
            }
        }
    }
    const float* m = l + 4;
    float mx = -3.4e38f;
    for (int j = sub; j < nmodes; j += 8) mx = fmaxf(mx, m[j]);
    mx = fmaxf(mx, __shfl_xor(mx, 1));
    mx = fmaxf(mx, __shfl_xor(mx, 2));
    mx = fmaxf(mx, __shfl_xor(mx, 4));
    float se = 0.0f;
    for (int j = sub; j < nmodes; j += 8) se += expf(m[j] - mx);
    se += __shfl_xor(se, 1);
    se += __shfl_xor(se, 2);
    se += __shfl_xor(se, 4);
    if (!live) return;
    // selection, one rank per round: the row's 8 lanes each take the best of
    // their remaining columns (value descending, lowest column first on ties),
    // the group keeps the best of the eight, and its owner removes it
    float* mw = sL + r * (n3p + 1) + 4;  // the row's mode logits (lane sub owns columns sub + 8c)
    const int nsel = min(kmode, nmodes);
    for (int rk = 0; rk < nsel; ++rk) {
        float bv = -INFINITY;
        int bj = 0x7FFFFFFF;
        for (int j = sub; j < nmodes; j += 8) {
            const float v = mw[j];
            if (v > bv) { bv = v; bj = j; }
        }
#pragma unroll
        for (int msk = 1; msk < 8; msk <<= 1) {
            const float ov = __shfl_xor(bv, msk);
            const int oj = __shfl_xor(bj, msk);
            if (ov > bv || (ov == bv && oj < bj)) { bv = ov; bj = oj; }
        }
        // no column left above -inf (or NaN logits): the rank stays empty
        const bool found = bj < nmodes;
        if (found && (bj & 7) == sub) mw[bj] = -INFINITY;
        if (sub == 0) {
            out_idx[o + 4 * kmax + rk] = found ? bj : -1;
            if (out_prob) out_prob[o + 4 * kmax + rk] = found ? expf(bv - mx) / se : 0.0f;
        }
    }
    for (int j = nmodes + sub; j < kmode; j += 8) {
        out_idx[o + 4 * kmax + j] = -1;
        if (out_prob) out_prob[o + 4 * kmax + j] = 0.0f;
    }
}

// Seed-head top-k of the listed rows for one partition count: the logits
// h2 W3 + b3 of the head's 1024 columns (the same matrix instructions in the
// same K order as the GEMM, so the same values as the separate kernels),
// then the k largest (value descending, lowest column first on ties), one
// wave per row.  16 rows per 256-thread workgroup (one 16-row tile of the
// matrix instruction, so each 512 KB head is read once per 16 rows); each
// wave computes 256 of the columns.  A row's logits sit in LDS until the
// wave has them in registers; that LDS row then holds the selection's
// histogram and candidate list, so two workgroups fit a CU.
constexpr int kSR = 16, kSeedCols = 1024, kSL = kSeedCols + 1;
constexpr int kSeedBuckets = 256;  // histogram of a row's logits over [min, max]
constexpr int kSeedCand = 192;     // logits in the selected buckets ranked directly
static_assert(kSeedBuckets + 2 * kSeedCand <= kSL, "selection scratch fits a logit row");

// Orders one wave's LDS accesses (the waves of k_seed_topk handle different
// numbers of rows, so a workgroup barrier cannot be used there).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(256) k_seed_topk(const int* __restrict__ list, const int* __restrict__ list_count,
                                                    int rows, const bf16* __restrict__ h2,
                                                    const bf16* __restrict__ W3t, const float* __restrict__ b3, int k,
                                                    int* __restrict__ out_idx, int out_stride, int out_off) {
    __shared__ __align__(16) bf16 sA[16 * kFS];
    __shared__ float sL[kSR * kSL];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int n = list ? min(*list_count, rows) : rows;
    const int r0 = blockIdx.x * kSR;
    if (r0 >= n) return;  // the whole workgroup
    for (int i = t; i < 16 * (kFH / 8); i += 256) {
        const int r = i / (kFH / 8), c = 8 * (i % (kFH / 8));
        bf16x8 v;
        for (int e = 0; e < 8; ++e) v[e] = (bf16)0.0f;
        if (r < kSR && r0 + r < n) {
            const int row = list ? list[r0 + r] : r0 + r;
            v = *reinterpret_cast<const bf16x8*>(h2 + (size_t)row * kFH + c);
        }
        *reinterpret_cast<bf16x8*>(sA + r * kFS + c) = v;
    }
    __syncthreads();
    {
        const int lr = lane & 15, lk = 8 * (lane >> 4);
        f32x4 acc[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int ks = 0; ks < kFH / 32; ++ks) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(sA + lr * kFS + 32 * ks + lk);
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const bf16x8 w =
                    *reinterpret_cast<const bf16x8*>(W3t + (size_t)(256 * wave + 16 * j + lr) * kFH + 32 * ks + lk);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc[j], 0, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int col = 256 * wave + 16 * j + lr;
            const float bias = b3[col];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 4 * (lane >> 4) + r;
                if (row < kSR) sL[row * kSL + col] = acc[j][r] + bias;
            }
        }
    }
    __syncthreads();
    for (int rr = wave; rr < kSR && r0 + rr < n; rr += 4) {
        const int row = list ? list[r0 + rr] : r0 + rr;
        float v[kSeedCols / 64];
#pragma unroll
        for (int m = 0; m < kSeedCols / 64; ++m) v[m] = sL[rr * kSL + lane + 64 * m];
        // Bucket threshold: the row's logits are histogrammed into 256
        // buckets over [min, max] (a monotone map, so a larger logit never
        // falls into a lower bucket); B is the highest bucket with at least k
        // logits in buckets >= B, so those buckets hold the k largest.  When
        // at most kSeedCand logits are there, they are ranked directly (value
        // descending, lowest column first).  Otherwise (heavy ties, or a
        // non-finite logit) the repeated arg-max below selects them.
        float lo = v[0], hi = v[0];
        bool finite = true;
#pragma unroll
        for (int m = 0; m < kSeedCols / 64; ++m) {
            lo = fminf(lo, v[m]);
            hi = fmaxf(hi, v[m]);
            finite = finite && isfinite(v[m]);
        }
#pragma unroll
        for (int msk = 32; msk > 0; msk >>= 1) {
            lo = fminf(lo, __shfl_xor(lo, msk));
            hi = fmaxf(hi, __shfl_xor(hi, msk));
        }
        finite = __ballot(!finite) == 0;
        int* s_h = reinterpret_cast<int*>(sL + rr * kSL);  // [kSeedBuckets], the row's logits are in v now
        float* s_cv = sL + rr * kSL + kSeedBuckets;       // [kSeedCand] candidate values
        int* s_ci = reinterpret_cast<int*>(s_cv + kSeedCand);  // [kSeedCand] candidate columns
        wave_sync();
        for (int i = lane; i < kSeedBuckets; i += 64) s_h[i] = 0;
        wave_sync();
        const float scale = (finite && hi > lo) ? (float)(kSeedBuckets - 1) / (hi - lo) : 0.0f;
        int bk[kSeedCols / 64];
#pragma unroll
        for (int m = 0; m < kSeedCols / 64; ++m) {
            bk[m] = min(kSeedBuckets - 1, max(0, (int)((v[m] - lo) * scale)));
            atomicAdd(&s_h[bk[m]], 1);
        }
        wave_sync();
        // suffix counts: lane l holds buckets 4l .. 4l + 3
        const int h0 = s_h[4 * lane], h1 = s_h[4 * lane + 1], h2c = s_h[4 * lane + 2], h3 = s_h[4 * lane + 3];
        int suf = h0 + h1 + h2c + h3;  // inclusive suffix over lanes lane .. 63
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_down(suf, d);
            if (lane + d < 64) suf += o;
        }
        const int s3 = suf - h0 - h1 - h2c, s2 = s3 + h2c, s1 = s2 + h1;
        int bsel = -1;
        if (s3 >= k) bsel = 4 * lane + 3;
        else if (s2 >= k) bsel = 4 * lane + 2;
        else if (s1 >= k) bsel = 4 * lane + 1;
        else if (suf >= k) bsel = 4 * lane;
#pragma unroll
        for (int msk = 32; msk > 0; msk >>= 1) bsel = max(bsel, __shfl_xor(bsel, msk));
        int mine = 0;
#pragma unroll
        for (int m = 0; m < kSeedCols / 64; ++m) mine += bk[m] >= bsel;
        int before = mine;  // inclusive prefix over the lanes
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(before, d);
            if (lane >= d) before += o;
        }
        const int total = __shfl(before, 63);
        if (finite && scale > 0.0f && total <= kSeedCand) {
            int slot = before - mine;
#pragma unroll
            for (int m = 0; m < kSeedCols / 64; ++m)
                if (bk[m] >= bsel) {
                    s_cv[slot] = v[m];
                    s_ci[slot] = lane + 64 * m;
                    ++slot;
                }
            wave_sync();
            for (int q = lane; q < total; q += 64) {
                const float cv = s_cv[q];
                const int ci = s_ci[q];
                int r = 0;
                for (int j = 0; j < total; ++j) {
                    const float o = s_cv[j];
                    r += (o > cv) || (o == cv && s_ci[j] < ci);
                }
                if (r < k) out_idx[(size_t)row * out_stride + out_off + r] = ci;
            }
            wave_sync();
            continue;
        }
        for (int j = 0; j < k; ++j) {
            float bv = -3.4e38f;
            int bi = 0x7FFFFFFF;
#pragma unroll
            for (int m = 0; m < kSeedCols / 64; ++m)
                if (v[m] > bv) { bv = v[m]; bi = lane + 64 * m; }
#pragma unroll
            for (int msk = 32; msk > 0; msk >>= 1) {
                const float ov = __shfl_xor(bv, msk);
                const int oi = __shfl_xor(bi, msk);
                if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
            }
            const bool found = bi < kSeedCols;
            if (found && (bi & 63) == lane) {
#pragma unroll
                for (int m = 0; m < kSeedCols / 64; ++m)
                    if (bi == lane + 64 * m) v[m] = -INFINITY;
            }
            if (lane == 0) out_idx[(size_t)row * out_stride + out_off + j] = found ? bi : -1;
        }
    }
}
}  // namespace

extern "C" {

astcg_status astcg_mlp_gemm(int m, int n, int k, const void* a, int lda, int trans_a, const void* b, int ldb,
                            int trans_b, float* c, int ldc, const float* bias, int accumulate, void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    if (m <= 0 || n <= 0 || k <= 0 || !a || !b || !c) return ASTCG_ERR_INVALID_ARGUMENT;
    const dim3 grid((n + kBN - 1) / kBN, (m + kBM - 1) / kBM);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bf16* A = static_cast<const bf16*>(a);
    const bf16* B = static_cast<const bf16*>(b);
    if (!trans_a && !trans_b) hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(256), 0, s, m, n, k, A, lda, B, ldb, c, ldc, bias, accumulate);
    else if (!trans_a && trans_b) hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(256), 0, s, m, n, k, A, lda, B, ldb, c, ldc, bias, accumulate);
    else if (trans_a && !trans_b) hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(256), 0, s, m, n, k, A, lda, B, ldb, c, ldc, bias, accumulate);
    else hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(256), 0, s, m, n, k, A, lda, B, ldb, c, ldc, bias, accumulate);
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}

astcg_status astcg_mlp_gelu(const float* z, void* h, long long n, void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    if (!z || !h || n < 0) return ASTCG_ERR_INVALID_ARGUMENT;
    if (n) hipLaunchKernelGGL(k_gelu_fwd, dim3(grid1(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), z,
                              static_cast<bf16*>(h), n);
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}

astcg_status astcg_mlp_gelu_backward(const float* dh, const float* z, void* dz, float* dz32, long long n,
                                     void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    if (!dh || !z || !dz || n < 0) return ASTCG_ERR_INVALID_ARGUMENT;
    if (n) hipLaunchKernelGGL(k_gelu_bwd, dim3(grid1(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), dh, z,
                              static_cast<bf16*>(dz), dz32, n);
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}

astcg_status astcg_mlp_to_bf16(const float* x, void* y, long long n, void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    if (!x || !y || n < 0) return ASTCG_ERR_INVALID_ARGUMENT;
    if (n) hipLaunchKernelGGL(k_to_bf16, dim3(grid1(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), x,
                              static_cast<bf16*>(y), n);
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}

astcg_status astcg_mlp_colsum(const float* x, int rows, int cols, float* out, int accumulate, void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    if (!x || !out || rows < 0 || cols <= 0) return ASTCG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(k_colsum, dim3(grid1(cols, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), x, rows, cols,
                       out, accumulate);
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}

astcg_status astcg_mlp_softmax_xent(const float* logits, int rows, int ld, int nheads, const int* head_off,
                                    const int* head_len, const int* labels, float scale, void* dlogits,
                                    float* dlogits32, float* loss, void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    if (!logits || rows < 0 || nheads <= 0 || !head_off || !head_len || !labels || !dlogits)
        return ASTCG_ERR_INVALID_ARGUMENT;
    if (rows)
        hipLaunchKernelGGL(k_softmax_xent, dim3(rows), dim3(64), 0, static_cast<hipStream_t>(stream), logits, rows, ld,
                           nheads, head_off, head_len, labels, scale, static_cast<bf16*>(dlogits), dlogits32, loss);
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}

astcg_status astcg_mlp_rank_xent(const float* logits, int rows, int ld, int ngroups, int ncand, const int* cols,
                                 const int* target, float scale, void* dlogits, float* dlogits32, float* loss,
                                 void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    if (!logits || rows < 0 || ngroups <= 0 || ncand <= 0 || ncand > 64 || !cols || !target || !dlogits)
        return ASTCG_ERR_INVALID_ARGUMENT;
    if (rows)
        hipLaunchKernelGGL(k_rank_xent, dim3(grid1(rows, 64)), dim3(64), 0, static_cast<hipStream_t>(stream), logits,
                           rows, ld, ngroups, ncand, cols, target, scale, static_cast<bf16*>(dlogits), dlogits32, loss);
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}

astcg_status astcg_mlp_adam(float* w, const float* g, float* m, float* v, void* w16, long long n, float lr,
                            float beta1, float beta2, float eps, int step, float weight_decay, void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    if (!w || !g || !m || !v || n < 0 || step < 1) return ASTCG_ERR_INVALID_ARGUMENT;
    float c1 = 1.0f, c2 = 1.0f;
    for (int i = 0; i < step && i < 100000; ++i) { c1 *= beta1; c2 *= beta2; }
    c1 = 1.0f - c1;
    c2 = 1.0f - c2;
    if (n)
        hipLaunchKernelGGL(k_adam, dim3(grid1(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), w, g, m, v,
                           static_cast<bf16*>(w16), n, lr, beta1, beta2, eps, c1, c2, weight_decay);
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}

astcg_status astcg_mlp_features(const uint8_t* d_rgba, int width, int height, int block_x, int block_y,
                                int first_block, int num_blocks, void* x, int ldx, void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    if (!d_rgba || !x || width <= 0 || height <= 0 || block_x < 4 || block_y < 4 || num_blocks < 0 ||
        ldx < 3 * block_x * block_y)
        return ASTCG_ERR_INVALID_ARGUMENT;
    const long long n = (long long)num_blocks * block_x * block_y;
    if (n)
        hipLaunchKernelGGL(k_features, dim3(grid1(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), d_rgba,
                           width, height, block_x, block_y, first_block, num_blocks, static_cast<bf16*>(x), ldx);
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}

astcg_status astcg_mlp_topk(const float* logits, int rows, int ld, int nheads, const int* head_off,
                            const int* head_len, int k, int* out_idx, float* out_prob, void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    return astcg_mlp_topk_heads(logits, rows, ld, nheads, head_off, head_len, k, nullptr, out_idx, out_prob, stream);
}

astcg_status astcg_mlp_topk_heads(const float* logits, int rows, int ld, int nheads, const int* head_off,
                                  const int* head_len, int kmax, const int* head_k, int* out_idx, float* out_prob,
                                  void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    if (!logits || rows < 0 || nheads <= 0 || !head_off || !head_len || kmax <= 0 || !out_idx)
        return ASTCG_ERR_INVALID_ARGUMENT;
    if (rows)
        hipLaunchKernelGGL(k_topk, dim3(rows, nheads), dim3(64), 0, static_cast<hipStream_t>(stream), logits, rows, ld,
                           nheads, head_off, head_len, kmax, head_k, out_idx, out_prob);
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}


astcg_status astcg_mlp_infer(int rows, const void* x, int ldx, int nin, const void* w1t, const float* b1,
                             const void* w2t, const float* b2, const void* w3t, const float* b3, int n3, int nmodes,
                             int kmax, int kpc, int kmode, int* out_idx, float* out_prob, void* h2_out, void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    const int kin = (nin + 63) / 64 * 64, n3p = (n3 + 15) / 16 * 16;
    if (!x || !w1t || !b1 || !w2t || !b2 || !w3t || !b3 || !out_idx || rows < 0 || nin <= 0 || kin > 256 ||
        ldx < nin || n3 < 4 + nmodes || nmodes <= 0 || n3p > 256 || kpc < 0 || kpc > 4 || kpc > kmax || kmode < 0 ||
        kmode > kmax)
        return ASTCG_ERR_INVALID_ARGUMENT;
    if (!rows) return ASTCG_OK;
    const size_t lds = 2 * sizeof(bf16) * kFR * kFS + sizeof(float) * kFR * (n3p + 1);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid(grid1(rows, kFR));
    const bf16 *X = static_cast<const bf16*>(x), *W1 = static_cast<const bf16*>(w1t), *W2 = static_cast<const bf16*>(w2t),
               *W3 = static_cast<const bf16*>(w3t);
    bf16* H2 = static_cast<bf16*>(h2_out);
#define ASTCG_INFER(KIN)                                                                                             \
    hipLaunchKernelGGL(k_infer<KIN>, grid, dim3(256), lds, s, rows, X, ldx, nin, W1, b1, W2, b2, W3, b3, n3p, nmodes, \
                       kmax, kpc, kmode, out_idx, out_prob, H2)
    switch (kin) {
        case 64: ASTCG_INFER(64); break;
        case 128: ASTCG_INFER(128); break;
        case 192: ASTCG_INFER(192); break;
        default: ASTCG_INFER(256); break;
    }
#undef ASTCG_INFER
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}

astcg_status astcg_mlp_seed_topk(const int* list, const int* list_count, int rows, const void* h2, const void* w3t,
                                 const float* b3, int k, int* out_idx, int out_stride, int out_off, void* stream) {
    (void)hipGetLastError();  // a stale error of an earlier call is not this call's
    if (!h2 || !w3t || !b3 || !out_idx || rows < 0 || k < 0 || k > kSeedCols || out_off < 0 ||
        out_stride < out_off + k || (list_count && !list) || (list && !list_count))
        return ASTCG_ERR_INVALID_ARGUMENT;
    if (!rows || !k) return ASTCG_OK;
    hipLaunchKernelGGL(k_seed_topk, dim3(grid1(rows, kSR)), dim3(256), 0, static_cast<hipStream_t>(stream), list,
                       list_count, rows, static_cast<const bf16*>(h2), static_cast<const bf16*>(w3t), b3, k, out_idx,
                       out_stride, out_off);
    return hipGetLastError() == hipSuccess ? ASTCG_OK : ASTCG_ERR_DEVICE;
}
}  // extern "C"
