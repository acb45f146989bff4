/*
 * astcg predictor kernels: the learned candidate policy's MLP building
 * blocks (gfx950 matrix cores, BF16 operands, FP32 accumulation).
 *
 * No counterpart in the reference (Jure-Z/ASTC_implementation has no learned
 * component); these entry points let a host driver train the candidate
 * predictor on labels produced by the encoder's own exhaustive search and run
 * it at encode time.  All pointers are device memory of the current HIP
 * device; work is enqueued on `stream` (a hipStream_t, NULL = null stream).
 * Matrices are row-major; "bf16" buffers hold IEEE bfloat16 values.
 */
#ifndef ASTCG_PREDICTOR_H
#define ASTCG_PREDICTOR_H

#include "astcg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* c[m x n] (fp32, row stride ldc) = op(a)[m x k] * op(b)[k x n] (+ bias[n])
 * (+ c when accumulate != 0).  a (bf16) is m x k with row stride lda, or k x m
 * when trans_a; b (bf16) is k x n with row stride ldb, or n x k when trans_b. */
astcg_status astcg_mlp_gemm(int m, int n, int k, const void* a, int lda, int trans_a, const void* b, int ldb,
                            int trans_b, float* c, int ldc, const float* bias, int accumulate, void* stream);

/* h (bf16) = gelu(z) elementwise, gelu(z) = z * Phi(z) (erf form). */
astcg_status astcg_mlp_gelu(const float* z, void* h, long long n, void* stream);

/* dz = dh * gelu'(z): bf16 into dz and, if dz32 is not NULL, fp32 into dz32. */
astcg_status astcg_mlp_gelu_backward(const float* dh, const float* z, void* dz, float* dz32, long long n,
                                     void* stream);

/* y (bf16) = x (fp32). */
astcg_status astcg_mlp_to_bf16(const float* x, void* y, long long n, void* stream);

/* out[c] (+)= sum over rows of x[r][c]. */
astcg_status astcg_mlp_colsum(const float* x, int rows, int cols, float* out, int accumulate, void* stream);

/* Softmax cross-entropy over `nheads` logit ranges per row (head h covers
 * columns head_off[h] .. head_off[h] + head_len[h] - 1; labels[r * nheads + h]
 * is the target column within the head, < 0 = no target).  Writes
 * dlogits = (softmax - onehot) * scale (bf16, and fp32 into dlogits32 if not
 * NULL) and loss[r] = the summed negative log-likelihoods (if not NULL). */
astcg_status astcg_mlp_softmax_xent(const float* logits, int rows, int ld, int nheads, const int* head_off,
                                    const int* head_len, const int* labels, float scale, void* dlogits,
                                    float* dlogits32, float* loss, void* stream);

/* Softmax cross-entropy restricted to candidate columns: group g of row r
 * (ngroups per row) is the ncand (<= 64) logit columns
 * cols[(r * ngroups + g) * ncand + j] (< 0 = absent), target[r * ngroups + g]
 * the index j of the correct one (< 0 = no target).  ADDS
 * (softmax - onehot) * scale at those columns into dlogits (bf16) and
 * dlogits32 (fp32, if not NULL), and the negative log-likelihoods into loss[r]
 * (if not NULL): a second loss on top of astcg_mlp_softmax_xent, used to rank
 * the compat partition positions with the seed heads. */
astcg_status astcg_mlp_rank_xent(const float* logits, int rows, int ld, int ngroups, int ncand, const int* cols,
                                 const int* target, float scale, void* dlogits, float* dlogits32, float* loss,
                                 void* stream);

/* Adam step `step` (>= 1, bias-corrected) with decoupled weight decay on fp32
 * master weights w; refreshes the bf16 copy w16 when not NULL. */
astcg_status astcg_mlp_adam(float* w, const float* g, float* m, float* v, void* w16, long long n, float lr,
                            float beta1, float beta2, float eps, int step, float weight_decay, void* stream);

/* Predictor input: for blocks first_block .. first_block + num_blocks - 1 of
 * an RGBA8 image (row-major blocks, edge texels clamped), writes the texels'
 * R, G, B / 255 as bf16 rows of 3 * block_x * block_y values (row stride ldx). */
astcg_status astcg_mlp_features(const uint8_t* d_rgba, int width, int height, int block_x, int block_y,
                                int first_block, int num_blocks, void* x, int ldx, void* stream);

/* The k best columns of every head of every row (value descending, lowest
 * column first on ties) into out_idx[(r * nheads + h) * k + j], and their
 * softmax probabilities into out_prob when not NULL. */
astcg_status astcg_mlp_topk(const float* logits, int rows, int ld, int nheads, const int* head_off,
                            const int* head_len, int k, int* out_idx, float* out_prob, void* stream);

/* astcg_mlp_topk with a per-head count: head h keeps head_k[h] (<= kmax,
 * device array; 0 skips the head) entries of its kmax-wide output slot. */
astcg_status astcg_mlp_topk_heads(const float* logits, int rows, int ld, int nheads, const int* head_off,
                                  const int* head_len, int kmax, const int* head_k, int* out_idx, float* out_prob,
                                  void* stream);

/* Fused inference for the encoder's search plan, hidden width 256: per row
 * of x (bf16, nin <= 256 inputs, row stride ldx) computes
 * h1 = gelu(x W1 + b1), h2 = gelu(h1 W2 + b2) (each rounded to bf16) and the
 * logits h2 W3 + b3 of n3 output columns (4 + nmodes <= n3 <= 256): columns
 * 0..3 are the partition-count head, 4 .. 4 + nmodes - 1 the trial-mode head.
 * Writes the softmax top-k of both heads (kpc <= min(4, kmax) and
 * kmode <= kmax entries, value descending, lowest column first on ties, -1
 * past the head's end) to out_idx / out_prob[(r * 5 + h) * kmax + j] for
 * h = 0 and h = 4 (out_prob may be NULL), and h2 to h2_out (bf16, row stride
 * 256) when not NULL.  Weights transposed and zero-padded, bf16: w1t
 * [256][nin rounded up to 64], w2t [256][256], w3t [n3 rounded up to 16][256];
 * b3 has n3 rounded up to 16 entries. */
astcg_status astcg_mlp_infer(int rows, const void* x, int ldx, int nin, const void* w1t, const float* b1,
                             const void* w2t, const float* b2, const void* w3t, const float* b3, int n3, int nmodes,
                             int kmax, int kpc, int kmode, int* out_idx, float* out_prob, void* h2_out, void* stream);

/* Seed-head top-k for one partition count, hidden width 256: for row r of
 * the list (list[0 .. min(*list_count, rows)), or rows 0 .. rows - 1 when
 * list is NULL), the k <= 1024 largest logits h2[row] W3 + b3 of the head's
 * 1024 columns (value descending, lowest column first on ties, -1 past the
 * head's end) to out_idx[row * out_stride + out_off + j], j < k; the same
 * values as astcg_mlp_gemm + astcg_mlp_topk_heads.  h2: bf16 [row][256];
 * w3t: the head's transposed bf16 weights [1024][256]; b3: [1024]. */
astcg_status astcg_mlp_seed_topk(const int* list, const int* list_count, int rows, const void* h2, const void* w3t,
                                 const float* b3, int k, int* out_idx, int out_stride, int out_off, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ASTCG_PREDICTOR_H */
