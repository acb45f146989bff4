// gfx950 kernels of the ASTC block search.
//
// Each kernel maps one work item of the reference's WGSL passes to one lane:
//   k_partition   passes 001-007  (one lane per image block, partition counts 2..4)
//   k_ideal_color passes 01, 08, 09, 10 (one lane per partitioned block)
//   k_dm_bm       passes 02-05 then 06, 07, 11 for every block mode of the
//                 decimation mode (one lane per partitioned block x decimation mode)
//   k_top         pass 12 (one lane per partitioned block)
//   k_refine      passes 13-17 x 6 (one lane per partitioned block x candidate);
//                 k_refine_q is its persistent work-queue form
//   k_pick        pass 18 + the host-side argmin of compression_pipeline.cpp:499-519
//   k_pack        symbolic_to_physical (physical_compression.cpp:843-1003)
// Every reduction the reference performs with float atomics is evaluated here
// sequentially in texel/weight order, which fixes the summation order.
// The file must be compiled with -ffp-contract=off.
#include <algorithm>
#include <mutex>
#include <vector>

#include "astcg_device.h"
#include "../../include/astcg_predictor.h"

namespace astcg {

__constant__ QuantTables c_qt;

void upload_quant_tables(hipStream_t s) {
    const QuantTables& qt = quant_tables();
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_qt), &qt, sizeof(QuantTables), 0, hipMemcpyHostToDevice, s);
}

namespace {

constexpr float kErrDefault = 1e37f;
constexpr int kFreeBits[4] = {111, 97, 94, 91};  // pass07:5
constexpr int kSteps[12] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32};
constexpr int kLevelsD[21] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256};

enum : int {
    FMT_LUMINANCE = 0, FMT_LUMINANCE_DELTA = 1, FMT_LUMINANCE_ALPHA = 4, FMT_LUMINANCE_ALPHA_DELTA = 5,
    FMT_RGB_SCALE = 6, FMT_RGB = 8, FMT_RGB_DELTA = 9, FMT_RGB_SCALE_ALPHA = 10, FMT_RGBA = 12, FMT_RGBA_DELTA = 13
};

__device__ __constant__ float kCW[4] = {0.30f * 2.25f, 0.59f * 2.25f, 0.11f * 2.25f, 1.0f};  // astc.h:60-63

// ---- WGSL operation semantics --------------------------------------------
__device__ __forceinline__ float wmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float wmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float wclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float wround(float x) { return rintf(x); }
__device__ __forceinline__ int32_t wf2i(float v) {
    if (!(v < 2147483520.0f)) return 2147483647;
    if (v < -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)v;
}
__device__ __forceinline__ uint32_t wf2u(float v) {
    if (!(v < 4294967040.0f)) return 0xFFFFFFFFu;
    if (v < 0.0f) return 0u;
    return (uint32_t)v;
}
__device__ __forceinline__ float dot4(const float* a, const float* b) {
    float s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s + a[3] * b[3];
}
__device__ __forceinline__ float dot3(const float* a, const float* b) {
    float s = a[0] * b[0];
    s = s + a[1] * b[1];
    return s + a[2] * b[2];
}
__device__ __forceinline__ void normalize4(const float* v, float* o) {
    const float l = sqrtf(dot4(v, v));
    o[0] = v[0] / l; o[1] = v[1] / l; o[2] = v[2] / l; o[3] = v[3] / l;
}
__device__ __forceinline__ void normalize3(const float* v, float* o) {
    const float l = sqrtf(dot3(v, v));
    o[0] = v[0] / l; o[1] = v[1] / l; o[2] = v[2] / l;
}

// Single-precision atan2 (Cephes atanf range reduction + polynomial); the
// same sequence of IEEE operations runs in the host checker.
__device__ float atanf_fixed(float x) {
    float sign = 1.0f;
    if (x < 0.0f) { sign = -1.0f; x = -x; }
    float y;
    if (x > 2.414213562373095f) { y = 1.5707963267948966f; x = -1.0f / x; }
    else if (x > 0.4142135623730950f) { y = 0.7853981633974483f; x = (x - 1.0f) / (x + 1.0f); }
    else y = 0.0f;
    const float z = x * x;
    float p = 8.05374449538e-2f;
    p = p * z - 1.38776856032e-1f;
    p = p * z + 1.99777106478e-1f;
    p = p * z - 3.33329491539e-1f;
    y = y + (p * z * x + x);
    return sign * y;
}
__device__ float atan2f_fixed(float y, float x) {
    if (x != x || y != y) return x + y;
    if (x == 0.0f) return y > 0.0f ? 1.5707963267948966f : (y < 0.0f ? -1.5707963267948966f : 0.0f);
    float w = 0.0f;
    if (x < 0.0f) w = (y < 0.0f) ? -3.141592653589793f : 3.141592653589793f;
    return w + atanf_fixed(y / x);
}

// ---- block access ----------------------------------------------------------
__device__ __forceinline__ TexDesc tex_of(const ImageArgs& A, int g) {
    // fields are selected one by one: a pointer that could point at either the
    // kernel argument or the table would force the arguments into scratch
    int lo = 0;
    if (A.ntex > 1) {
        int hi = A.ntex - 1;  // last texture starting at or before block g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (A.tex[mid].first_block <= g) lo = mid;
            else hi = mid - 1;
        }
    }
    const bool one = A.ntex == 1;
    const TexDesc* p = A.tex + lo;
    TexDesc t;
    t.rgba = one ? A.t0.rgba : p->rgba;
    t.out = one ? A.t0.out : p->out;
    t.width = one ? A.t0.width : p->width;
    t.height = one ? A.t0.height : p->height;
    t.blocks_x = one ? A.t0.blocks_x : p->blocks_x;
    t.first_block = one ? A.t0.first_block : p->first_block;
    t.num_blocks = one ? A.t0.num_blocks : p->num_blocks;
    t.pad = 0;
    return t;
}

// Texel origin of one block of a texture (edge blocks clamp to the last
// row/column, compression_pipeline.cpp:33,38).
struct BlockLoc {
    const uint8_t* rgba;
    int width, height;
    int x0, y0;
    int local;  // block index within its texture
    int tex_blocks;
};

__device__ __forceinline__ BlockLoc block_loc(const ImageArgs& A, const DevTables& T, int g) {
    const TexDesc t = tex_of(A, g);
    BlockLoc l;
    l.rgba = t.rgba;
    l.width = t.width;
    l.height = t.height;
    l.local = g - t.first_block;
    l.tex_blocks = t.num_blocks;
    const int byi = l.local / t.blocks_x, bxi = l.local - byi * t.blocks_x;
    l.x0 = bxi * T.bx;
    l.y0 = byi * T.by;
    return l;
}

__device__ __forceinline__ uint32_t texel_rgba8(const BlockLoc& l, const DevTables& T, int i) {
    const int dy = i / T.bx, dx = i - dy * T.bx;
    const int x = min(l.x0 + dx, l.width - 1), y = min(l.y0 + dy, l.height - 1);
    return *reinterpret_cast<const uint32_t*>(l.rgba + ((size_t)y * l.width + x) * 4);
}

// (u8 / 255) * 65536 as IEEE float division and scaling round it, without a
// divide or a table: 65536 / 255 = 257 + 1 / 255, and 257 * u8 is an exact
// float, so one fused multiply-add with the rounded reciprocal of 255,
// fma(u8, RN(1/255), 257 * u8), rounds the exact 257 u8 + u8 / 255 once.
// Its error term u8 * (RN(1/255) - 1/255) is below 2^-33 * 255, which moves
// none of the 256 values across a rounding boundary (checked exhaustively in
// exact rational arithmetic; the parity tests compare with the oracle's
// division).
__device__ __forceinline__ float u8_search_value(uint32_t v) {
    const float x = (float)v;
    return __builtin_fmaf(x, 1.0f / 255.0f, x * 257.0f);
}

// a texel's channels as search values (arithmetic, not a table: four fewer
// dependent loads per texel)
__device__ __forceinline__ void texel_px(const BlockLoc& l, const DevTables& T, int i, float* o) {
    const uint32_t v = texel_rgba8(l, T, i);
    o[0] = u8_search_value(v & 0xFF);
    o[1] = u8_search_value((v >> 8) & 0xFF);
    o[2] = u8_search_value((v >> 16) & 0xFF);
    o[3] = u8_search_value(v >> 24);
}

// The chunk's block texels as one contiguous row per block ([block][texel],
// edge-clamped), so the kernels that walk a partitioned block's texels many
// times read a few cache lines with no coordinate arithmetic.  A thread per
// texel.
__global__ void __launch_bounds__(256) k_block_texels(ImageArgs A, DevTables T, uint32_t* out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int tc = T.tc;
    if (e >= (size_t)A.nb * tc) return;
    const int k = (int)(e / tc), t = (int)(e - (size_t)k * tc);
    out[e] = texel_rgba8(block_loc(A, T, A.blk0 + k), T, t);
}

struct Pb {
    int g;          // block index of the launch
    BlockLoc loc;
    int pc;         // partition count
    int entry;      // packed partition table entry (pc > 1)
    const uint8_t* tp;  // texel partitions (nullptr: all 0)
    int ppc[4];
    int seed;
    int pb_in_batch;
};

__device__ __forceinline__ int texel_part(const Pb& b, int i) { return b.tp ? b.tp[i] : 0; }

// Partitioned block pb of partition count pc -> chunk block k, candidate
// position c (0..3) and its slot in pb_part ([block][4]).
__device__ __forceinline__ void pb_decode(const ImageArgs& A, int pc, int pb, int* k, int* c, int* slot) {
    if (pc == 1) {
        *k = pb; *c = 0; *slot = pb;
    } else if (A.pb_list) {
        const int kc = A.pb_list[pb];
        *k = kc >> 2; *c = kc & 3; *slot = kc;
    } else {
        *k = pb >> 2; *c = pb & 3; *slot = pb;
    }
}

// Partitioned blocks in this launch (the learned policy's list length, or
// the launch's full count).
__device__ __forceinline__ int pb_limit(const ImageArgs& A, int pc, int npb) {
    return (pc > 1 && A.pb_count) ? *A.pb_count : npb;
}

__device__ __forceinline__ int pb_slot_of(const ImageArgs& A, int pc, int pb) {
    int k, c, slot;
    pb_decode(A, pc, pb, &k, &c, &slot);
    return slot;
}

__device__ Pb make_pb(const ImageArgs& A, const DevTables& T, const int16_t* pb_part, int pc, int pb) {
    Pb b;
    int k, c, slot;
    pb_decode(A, pc, pb, &k, &c, &slot);
    b.g = A.blk0 + k;
    b.loc = block_loc(A, T, b.g);
    b.pc = pc;
    // the reference encodes each texture in batches of 920 blocks
    const int j = b.loc.local % kRefBatch;
    b.pb_in_batch = pc == 1 ? j : j * 4 + c;
    if (pc == 1) {
        b.entry = 0;
        b.tp = nullptr;
        b.ppc[0] = T.tc; b.ppc[1] = b.ppc[2] = b.ppc[3] = 0;
        b.seed = 0;
    } else {
        const int e = pb_part[slot];
        b.entry = e;
        const size_t row = (size_t)(pc - 2) * kPartitionings + e;
        b.tp = T.part_of_texel + row * T.tc;
        for (int i = 0; i < 4; ++i) b.ppc[i] = T.part_counts[row * 4 + i];
        b.seed = T.part_seed[row];
    }
    return b;
}

__device__ __forceinline__ uint16_t pack_fmt4(const int* f) {
    return (uint16_t)((f[0] & 0xF) | ((f[1] & 0xF) << 4) | ((f[2] & 0xF) << 8) | ((f[3] & 0xF) << 12));
}
__device__ __forceinline__ int fmt_of(uint16_t v, int p) { return (v >> (4 * p)) & 0xF; }

// ---- passes 001-007: compat partition candidates ---------------------------
// A 64-lane workgroup handles 4 blocks with 16 lanes each.  The blocks'
// texels sit in LDS.  k-means runs texel-parallel: every cross-texel
// reduction in passes 001-003 is an integer sum, an integer max or a
// highest-index pick, so splitting it over lanes gives the same result as
// the sequential loop.  The 128 ordered candidates of pass006 are spread over
// the 16 lanes, and the in-order top-4 insertion of pass007 runs on the
// group's first lane.
constexpr int kPartBlocks = 4;
constexpr int kPartLanes = 16;
constexpr int kPartMaxR = (kMaxTexels + kPartLanes - 1) / kPartLanes;

__device__ __forceinline__ float dist_sq(const float* a, const float* b) {
    float d[4];
    for (int c = 0; c < 4; ++c) { const float t = a[c] - b[c]; d[c] = t * t; }
    return dot4(d, kCW);
}
__device__ __forceinline__ void ld4(const float4* p, int i, float* o) {
    const float4 v = p[i];
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
__device__ __forceinline__ uint32_t grp_max_u32(uint32_t v) {
    for (int m = 1; m < kPartLanes; m <<= 1) v = max(v, (uint32_t)__shfl_xor((int)v, m, kPartLanes));
    return v;
}
__device__ __forceinline__ int grp_max_i32(int v) {
    for (int m = 1; m < kPartLanes; m <<= 1) v = max(v, __shfl_xor(v, m, kPartLanes));
    return v;
}
__device__ __forceinline__ uint32_t grp_sum_u32(uint32_t v) {
    for (int m = 1; m < kPartLanes; m <<= 1) v += (uint32_t)__shfl_xor((int)v, m, kPartLanes);
    return v;
}
__device__ __forceinline__ uint64_t grp_or_u64(uint64_t v) {
    for (int m = 1; m < kPartLanes; m <<= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, kPartLanes);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, kPartLanes);
        v |= ((uint64_t)hi << 32) | lo;
    }
    return v;
}

// pass001-003: texel -> cluster after the reference's 4 k-means iterations;
// lane `sub` of the group owns texels sub, sub + 16, ...
__device__ __forceinline__ void kmeans_assign_grp(const float4* kpx, int tc, int pc, int sub, uint8_t* asg) {
    float centers[4][4] = {};
    ld4(kpx, 145897 % tc, centers[0]);
    float dist[kPartMaxR];
#pragma unroll
    for (int r = 0; r < kPartMaxR; ++r) {
        const int t = sub + r * kPartLanes;
        dist[r] = 0.0f;
        if (t < tc) { float px[4]; ld4(kpx, t, px); dist[r] = dist_sq(centers[0], px); }
    }
#pragma unroll
    for (int p = 1; p < 4; ++p) {
        if (p >= pc) continue;
        uint32_t mx = 0;
#pragma unroll
        for (int r = 0; r < kPartMaxR; ++r)
            if (sub + r * kPartLanes < tc) mx = max(mx, __float_as_uint(dist[r]));
        mx = grp_max_u32(mx);
        int idx = 0;  // the highest texel holding the maximum
#pragma unroll
        for (int r = 0; r < kPartMaxR; ++r) {
            const int t = sub + r * kPartLanes;
            if (t < tc && __float_as_uint(dist[r]) == mx) idx = t;
        }
        idx = grp_max_i32(idx);
        ld4(kpx, idx, centers[p]);
        if (p < pc - 1) {
#pragma unroll
            for (int r = 0; r < kPartMaxR; ++r) {
                const int t = sub + r * kPartLanes;
                if (t < tc) { float px[4]; ld4(kpx, t, px); dist[r] = wmin(dist[r], dist_sq(centers[p], px)); }
            }
        }
    }
    for (int it = 0; it < 4; ++it) {
        uint32_t counts[4] = {0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < kPartMaxR; ++r) {
            const int t = sub + r * kPartLanes;
            if (t >= tc) continue;
            float px[4];
            ld4(kpx, t, px);
            float bd = 1e30f;
            int bp = 0;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                if (p >= pc) continue;
                const float d = dist_sq(px, centers[p]);
                if (d < bd) { bd = d; bp = p; }
            }
            asg[t] = (uint8_t)bp;
#pragma unroll
            for (int p = 0; p < 4; ++p) counts[p] += bp == p ? 1u : 0u;
        }
        bool problem = false;
#pragma unroll
        for (int p = 0; p < 4; ++p)
            if (p < pc) problem |= grp_sum_u32(counts[p]) == 0;
        if (problem && sub < pc) asg[sub] = (uint8_t)sub;  // texels 0..pc-1 (pc <= 4 < 16)
        uint32_t sums[4][4] = {}, cnt[4] = {};
#pragma unroll
        for (int r = 0; r < kPartMaxR; ++r) {
            const int t = sub + r * kPartLanes;
            if (t >= tc) continue;
            float px[4];
            ld4(kpx, t, px);
            const int p = asg[t];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t on = p == q ? 1u : 0u;
                for (int c = 0; c < 4; ++c) sums[q][c] += on * wf2u(px[c]);
                cnt[q] += on;
            }
        }
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            if (p >= pc) continue;
            const uint32_t n = grp_sum_u32(cnt[p]);
            for (int c = 0; c < 4; ++c) {
                const uint32_t s = grp_sum_u32(sums[p][c]);
                centers[p][c] = n > 0 ? (float)s / (float)n : 0.0f;
            }
        }
    }
}

// pass004: minimum-over-permutations coverage mismatch for one partitioning;
// lanes OR their share of the 64 k-means texels into the group's bitmaps
__device__ __forceinline__ uint32_t mismatch_count_grp(const DevTables& T, const uint8_t* asg, int pc, int entry, int sub) {
    uint64_t a[4] = {0, 0, 0, 0};
    for (int l = sub; l < 64; l += kPartLanes) {
        const int p = asg[T.kmeans_texels[l]];
        for (int q = 0; q < 4; ++q) a[q] |= (p == q) ? (1ull << l) : 0ull;
    }
    for (int q = 0; q < 4; ++q) a[q] = grp_or_u64(a[q]);
    const uint64_t* bb = T.coverage + ((size_t)(pc - 2) * kPartitionings + entry) * 4;
    auto P = [&](int i, int j) { return __popcll(a[i] ^ bb[j]); };
    if (pc == 2) return (uint32_t)(min(P(0, 0) + P(1, 1), P(0, 1) + P(1, 0)) / 2);
    if (pc == 3) {
        const int v0 = min(P(1, 1) + P(2, 2), P(1, 2) + P(2, 1)) + P(0, 0);
        const int v1 = min(P(1, 0) + P(2, 2), P(1, 2) + P(2, 0)) + P(0, 1);
        const int v2 = min(P(1, 0) + P(2, 1), P(1, 1) + P(2, 0)) + P(0, 2);
        return (uint32_t)(min(v0, min(v1, v2)) / 2);
    }
    const int mx23 = min(P(2, 2) + P(3, 3), P(2, 3) + P(3, 2));
    const int mx13 = min(P(2, 1) + P(3, 3), P(2, 3) + P(3, 1));
    const int mx12 = min(P(2, 1) + P(3, 2), P(2, 2) + P(3, 1));
    const int mx03 = min(P(2, 0) + P(3, 3), P(2, 3) + P(3, 0));
    const int mx02 = min(P(2, 0) + P(3, 2), P(2, 2) + P(3, 0));
    const int mx01 = min(P(2, 1) + P(3, 0), P(2, 0) + P(3, 1));
    const int v0 = P(0, 0) + min(P(1, 1) + mx23, min(P(1, 2) + mx13, P(1, 3) + mx12));
    const int v1 = P(0, 1) + min(P(1, 0) + mx23, min(P(1, 2) + mx03, P(1, 3) + mx02));
    const int v2 = P(0, 2) + min(P(1, 1) + mx03, min(P(1, 0) + mx13, P(1, 3) + mx01));
    const int v3 = P(0, 3) + min(P(1, 1) + mx02, min(P(1, 2) + mx01, P(1, 0) + mx12));
    return (uint32_t)(min(v0, min(v1, min(v2, v3))) / 2);
}

// pass006: line-fit errors of one candidate partitioning, with the
// reference's per-invocation avg[]/dir[] arrays (only lane p holds
// partition p's average and direction).  Its cross-texel sums are integer.
__device__ __forceinline__ void eval_candidate(const float4* bpx, const DevTables& T, int pc, int entry, float* eu, float* es) {
    // Per-partition state is indexed through compile-time partition numbers
    // (selects), so it stays in registers.  Only texels 0..pc-1 can be
    // "own" texels (own: partition == texel index), so those are peeled.
    const int tc = T.tc;
    const size_t row = (size_t)(pc - 2) * kPartitionings + entry;
    const uint8_t* pot = T.part_of_texel + row * tc;
    uint32_t sums[4][4] = {}, cnt[4] = {};
    for (int t = 0; t < tc; ++t) {
        float px[4];
        ld4(bpx, t, px);
        const int p = pot[t];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < pc && p == q) {
                cnt[q]++;
                for (int c = 0; c < 4; ++c) sums[q][c] += wf2u(px[c]);
            }
    }
    float avg[4][4] = {};
    bool has_avg[4] = {false, false, false, false};
#pragma unroll
    for (int p = 0; p < 4; ++p)
        if (p < pc && (float)cnt[p] > 0.0f) {
            const float n = (float)cnt[p];
            for (int c = 0; c < 4; ++c) avg[p][c] = (float)sums[p][c] / n;
            has_avg[p] = true;
        }
    int32_t axis[4][4][4] = {};
    auto axis_texel = [&](int t, bool own_ok, const float* own_avg) __attribute__((always_inline)) {
        float px[4];
        ld4(bpx, t, px);
        const int p = pot[t];
        const bool own = own_ok && p == t;
        float datum[4];
        for (int c = 0; c < 4; ++c) datum[c] = px[c] - (own ? own_avg[c] : 0.0f);
        int32_t di[4];
        for (int c = 0; c < 4; ++c) di[c] = wf2i(datum[c]);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < pc && p == q)
                for (int ax = 0; ax < 4; ++ax)
                    if (datum[ax] > 0.0f)
                        for (int c = 0; c < 4; ++c) axis[q][ax][c] = (int32_t)((uint32_t)axis[q][ax][c] + (uint32_t)di[c]);
    };
#pragma unroll
    for (int t = 0; t < 4; ++t)
        if (t < tc) axis_texel(t, t < pc && has_avg[t], avg[t]);
    for (int t = 4; t < tc; ++t) axis_texel(t, false, avg[0]);
    float dir[4][4] = {};
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        if (p >= pc) continue;
        float s[4][4];
        for (int ax = 0; ax < 4; ++ax)
            for (int c = 0; c < 4; ++c) s[ax][c] = (float)axis[p][ax][c];
        float best[4];
        for (int c = 0; c < 4; ++c) best[c] = s[0][c];
        float mp = dot4(s[0], s[0]);
        for (int ax = 1; ax < 4; ++ax) {
            const float pr = dot4(s[ax], s[ax]);
            if (pr > mp) { mp = pr; for (int c = 0; c < 4; ++c) best[c] = s[ax][c]; }
        }
        normalize4(best, dir[p]);
    }
    uint32_t usum = 0, ssum = 0;
    int32_t lmin[4] = {2147483647, 2147483647, 2147483647, 2147483647};
    int32_t lmax[4] = {-2147483647, -2147483647, -2147483647, -2147483647};
    auto line_texel = [&](int t, bool own_ok, const float* own_dir, const float* own_avg) __attribute__((always_inline)) {
        float px[4];
        ld4(bpx, t, px);
        const int p = pot[t];
        const bool own = own_ok && p == t;
        float ud[4], av[4];
        for (int c = 0; c < 4; ++c) {
            ud[c] = own ? own_dir[c] : 0.0f;
            av[c] = own ? own_avg[c] : 0.0f;
        }
        const float uparam = dot4(px, ud);
        float dm[4];
        for (int c = 0; c < 4; ++c) dm[c] = px[c] - av[c];
        const float s = dot4(dm, ud);
        float diff[4];
        for (int c = 0; c < 4; ++c) { const float d = px[c] - (av[c] + ud[c] * s); diff[c] = d * d; }
        const float udist = dot4(diff, kCW);
        float sd[4];
        normalize4(av, sd);
        const float ss = dot4(px, sd);
        for (int c = 0; c < 4; ++c) { const float d = px[c] - sd[c] * ss; diff[c] = d * d; }
        const float sdist = dot4(diff, kCW);
        usum += wf2u(udist);
        ssum += wf2u(sdist);
        const int32_t pi = wf2i(uparam * 4096.0f);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < pc && p == q) { lmin[q] = min(lmin[q], pi); lmax[q] = max(lmax[q], pi); }
    };
    // own texel: avg is used only when the partition has texels (it always
    // does when p == t), dir always
    float zero4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 4; ++t)
        if (t < tc) {
            float oa[4];
            for (int c = 0; c < 4; ++c) oa[c] = has_avg[t] ? avg[t][c] : 0.0f;
            line_texel(t, t < pc, dir[t], oa);
        }
    for (int t = 4; t < tc; ++t) line_texel(t, false, zero4, zero4);
    float tu = (float)usum, ts = (float)ssum;
    float wie = 0.055f;
    if (tc <= 20) wie = 0.03f;
    else if (tc <= 31) wie = 0.04f;
    else if (tc <= 41) wie = 0.05f;
    wie = wie * wie;
    const uint8_t* cnts = T.part_counts + row * 4;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        if (p >= pc) continue;
        const float minp = (float)lmin[p] / 4096.0f, maxp = (float)lmax[p] / 4096.0f;
        const float len = wmax(maxp - minp, 1e-7f);
        const float ew = (float)cnts[p] * wie;
        float a0[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (p == 0 && has_avg[0]) for (int c = 0; c < 4; ++c) a0[c] = avg[0][c];
        tu = tu + ew * (len * len);
        ts = ts + ew * dot4(a0, a0);
    }
    *eu = tu;
    *es = ts;
}

size_t partition_lds_bytes(const DevTables& T) {
    const size_t stride = T.tc + 1;
    return 2 * kPartBlocks * stride * sizeof(float4) + kPartBlocks * 128 * sizeof(float2) +
           kPartBlocks * (((size_t)T.tc + 3) & ~size_t(3));
}

// The learned policy's per-block decisions for partition count pc, shared by
// k_partition (which skips the blocks whose positions are not searched) and
// k_plan (which builds the lists); k is the block within the chunk, g its
// index in the launch chain.
__device__ __forceinline__ bool plan_early_out(const SearchPlan& P, int g) {
    // the block's best encoding so far is below the error limit
    return P.error_limit > 0.0f && P.best[g].error < P.error_limit;
}

__device__ __forceinline__ bool plan_hard(const SearchPlan& P, int pc, int k, int g) {
    // a block whose best 1-partition error is above full_limit gets the full
    // search from 2 partitions on
    return pc == 2 && P.full_limit > 0.0f && !P.fallback[k] && P.best[g].error > P.full_limit;
}

__device__ bool plan_keeps_count(const SearchPlan& P, int pc, int k, int g) {
    // pc is the count head's most likely count, or at least count_below
    // probable, or the block's best error so far is above count_limit
    if (P.count_limit > 0.0f && P.best[g].error > P.count_limit) return true;
    const int* ti = P.topk_idx + (size_t)k * 5 * P.kmax;
    const float* tp = P.topk_prob + (size_t)k * 5 * P.kmax;
    float pp = 0.0f;
    bool first = false;
    for (int j = 0; j < P.kmax && j < 4; ++j)
        if (ti[j] == pc - 1) { pp = tp[j]; first = j == 0; }
    return first || pp >= P.count_below;
}

// Whether k_plan searches block k at count pc from the compat positions.
__device__ bool plan_uses_positions(const SearchPlan& P, int pc, int k, int g) {
    if (plan_early_out(P, g)) return false;
    if (P.fallback[k] || plan_hard(P, pc, k, g)) return true;
    return P.seeds == 0 && plan_keeps_count(P, pc, k, g);
}

// Whether block k searches count pc from its predicted seeds (with more than
// 4 seeds kept per count, ranked by their line fits).
__device__ bool plan_uses_seeds(const SearchPlan& P, int pc, int k, int g) {
    if (P.seeds <= 0 || plan_early_out(P, g) || P.fallback[k] || plan_hard(P, pc, k, g)) return false;
    return plan_keeps_count(P, pc, k, g);
}

// Specialised per partition count (launched once per count), so the
// per-partition selects of the k-means and line-fit loops only cover the
// partitions that exist.
template <int PC>
__global__ void __launch_bounds__(64) k_partition(ImageArgs A, DevTables T, int16_t* pb_part, SearchPlan P,
                                                  int use_plan) {
    constexpr int pc = PC;
    extern __shared__ __align__(16) uint8_t smem[];
    const int tc = T.tc, lane = threadIdx.x, grp = lane / kPartLanes, sub = lane % kPartLanes;
    const int stride = tc + 1;  // float4 units; the pad spreads the 4 blocks over LDS banks
    float4* s_px = reinterpret_cast<float4*>(smem);        // blocks being partitioned
    float4* s_kpx = s_px + kPartBlocks * stride;           // blocks whose k-means clusters they read
    float2* s_err = reinterpret_cast<float2*>(s_kpx + kPartBlocks * stride);
    uint8_t* s_asg = reinterpret_cast<uint8_t*>(s_err + kPartBlocks * 128);
    const int tcp = (tc + 3) & ~3;
    const int k = blockIdx.x * kPartBlocks + grp;
    const int g = A.blk0 + k;
    const BlockLoc loc = block_loc(A, T, k < A.nb ? g : A.blk0);
    const int j = loc.local % kRefBatch, base = loc.local - j;  // within the texture
    const int B = min(kRefBatch, loc.tex_blocks - base);
    // partition_ordering and the partitioning errors are allocated 920*128
    // entries but indexed with stride 512, so in-batch blocks >= 230 read
    // zeros and keep table positions 0..3.
    // (learned policy: a block that is not searched at this count from the
    // compat positions needs no candidates)
    const bool work = k < A.nb && j * 512 + 127 < kRefBatch * 128 && (!use_plan || plan_uses_positions(P, pc, k, g));
    // pass004 writes slot block*1024*i: the score this block reads is
    // mismatch(j, partitioning 1), or mismatch(B-1, partitioning 0) for j == 0
    const int src = j == 0 ? g + B - 1 : g;
    if (work) {
        for (int t = sub; t < tc; t += kPartLanes) {
            float px[4];
            texel_px(loc, T, t, px);
            s_px[grp * stride + t] = make_float4(px[0], px[1], px[2], px[3]);
            if (src != g) texel_px(block_loc(A, T, src), T, t, px);
            s_kpx[grp * stride + t] = make_float4(px[0], px[1], px[2], px[3]);
        }
    }
    __syncthreads();
    uint8_t* asg = s_asg + grp * tcp;
    if (work) kmeans_assign_grp(s_kpx + grp * stride, tc, pc, sub, asg);
    __syncthreads();
    if (work) {
        const uint32_t v = mismatch_count_grp(T, asg, pc, j == 0 ? 0 : 1, sub);
        const uint16_t* ord = T.ordering + ((pc - 2) * 2 + (v > 0 ? 1 : 0)) * 128;
        for (int i = sub; i < 128; i += kPartLanes) {
            float eu, es;
            eval_candidate(s_px + grp * stride, T, pc, ord[i], &eu, &es);
            s_err[grp * 128 + i] = make_float2(eu, es);
        }
    }
    __syncthreads();
    if (k >= A.nb || sub != 0) return;
    int pos[4] = {0, 1, 2, 3};
    if (work) {
        int bui[4] = {0, 0, 0, 0}, bsi[4] = {0, 0, 0, 0};
        float bue[4] = {1e30f, 1e30f, 1e30f, 1e30f}, bse[4] = {1e30f, 1e30f, 1e30f, 1e30f};
        for (int i = 0; i < 128; ++i) {
            const float2 e = s_err[grp * 128 + i];
            const float eu = e.x, es = e.y;
            if (eu < bue[3]) {
                bue[3] = eu; bui[3] = i;
                for (int q = 3; q > 0; --q)
                    if (bue[q] < bue[q - 1]) { float t = bue[q]; bue[q] = bue[q - 1]; bue[q - 1] = t; int u = bui[q]; bui[q] = bui[q - 1]; bui[q - 1] = u; }
            }
            if (es < bse[3]) {
                bse[3] = es; bsi[3] = i;
                for (int q = 3; q > 0; --q)
                    if (bse[q] < bse[q - 1]) { float t = bse[q]; bse[q] = bse[q - 1]; bse[q - 1] = t; int u = bsi[q]; bsi[q] = bsi[q - 1]; bsi[q - 1] = u; }
            }
        }
        int emitted = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int idx = (i & 1) ? bsi[i >> 1] : bui[i >> 1];
            bool dup = emitted >= 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) dup |= q < emitted && pos[q] == idx;
            if (!dup) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q == emitted) pos[q] = idx;
                ++emitted;
            }
        }
        // pass007 uses the candidate position itself as the packed table index
    }
    for (int c = 0; c < 4; ++c) pb_part[k * 4 + c] = (int16_t)pos[c];
}

// ---- exhaustive partition candidates (code/best_partitionings.cpp:498-737) ----
// The CPU partition search the reference leaves commented out
// (code/compression_pipeline.cpp:363-389), run over every selected
// partitioning: weighted-random k-means seeding and three k-means passes,
// coverage mismatch of every selected partitioning, a stable counting sort by
// mismatch, the uncorrelated / same-chroma line fits of each partitioning in
// that order, the 4 best of each kept by the reference's insertion, then
// interleaved and de-duplicated.  One lane per block: every step is a
// per-block sequential loop whose float operations keep the reference's
// order.  The block's RGBA8 texels and k-means assignment sit in LDS
// ([texel][lane]); the mismatch counts and the ordering go to a per-block
// scratch column (the mode-error buffer of the same partition count, which is
// written only later), [row][block] so that the lanes' accesses coalesce.
__device__ __constant__ float kKmeansCutoffs[9] = {0.626220f, 0.932770f, 0.275454f, 0.318558f, 0.240113f,
                                                   0.009190f, 0.347661f, 0.731960f, 0.156391f};

__device__ __forceinline__ float ex_dist(const float* a, const float* b) {
    const float d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2], d3 = a[3] - b[3];
    return (((d0 * d0 * kCW[0]) + (d1 * d1 * kCW[1])) + (d2 * d2 * kCW[2])) + (d3 * d3 * kCW[3]);
}

__device__ __forceinline__ void ex_normalize(const float* v, float* o) {
    float n = ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3];
    if (n != 0.0f) {
        n = sqrtf(n);
        for (int c = 0; c < 4; ++c) o[c] = v[c] / n;
    } else {
        for (int c = 0; c < 4; ++c) o[c] = 0.0f;
    }
}

// insert_result (best_partitionings.cpp:460-496) into a sorted list of 4:
// the new entry goes before the first entry it does not exceed; written with
// compile-time indices only (the lists stay in registers)
__device__ __forceinline__ void ex_insert(float e, int id, float* be, int* bi) {
    if (e >= be[3]) return;
    int at = 4;
#pragma unroll
    for (int i = 3; i >= 0; --i)
        if (!(e > be[i])) at = i;
#pragma unroll
    for (int j = 3; j > 0; --j)
        if (j > at) { be[j] = be[j - 1]; bi[j] = bi[j - 1]; }
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (j == at) { be[j] = e; bi[j] = id; }
}

// The uncorrelated / same-chroma line fits of one partitioning (packed entry
// i of count pc) of a block (find_best_partition_candidates,
// compute_avgs_and_dirs_{4_comp,3_comp_rgb}, compute_error_squared_{rgba,rgb},
// best_partitionings.cpp:520-737): errors *ue_out, *se_out.  PX(t, px)
// gives texel t's search values.
template <int PC, int kExWords, class PXF>
__device__ __forceinline__ void ex_line_fit(const PXF& PX, const DevTables& T, int pc, int i, bool rgba, float wie,
                                            float* ue_out, float* se_out) {
    const int tc = T.tc;
    const size_t row = (size_t)(pc - 2) * kPartitionings + i;
    const uint8_t* pot = T.part_of_texel + row * tc;
    const uint8_t* cnts = T.part_counts + row * 4;
    // per-partition sums in texel order, and each partition's texels as
    // bit masks: the later passes visit only a partition's own texels,
    // in increasing order like the reference's texels_of_partition
    float sum[PC][4] = {};
    uint64_t mk[PC][kExWords] = {};
    for (int t = 0; t < tc; ++t) {
        float px[4];
        PX(t, px);
        const int p = pot[t];
        const uint64_t bit = 1ull << (t & 63);
#pragma unroll
        for (int q = 0; q < PC; ++q)
            if (p == q) {
                for (int c = 0; c < 4; ++c) sum[q][c] += px[c];
#pragma unroll
                for (int wd = 0; wd < kExWords; ++wd) mk[q][wd] |= (t >> 6) == wd ? bit : 0ull;
            }
    }
    float ue = 0.0f, se = 0.0f, len[PC], bdir[PC][4], bavg[PC][4];
#pragma unroll
    for (int p = 0; p < PC; ++p) {
        float avg[4];
        for (int c = 0; c < 4; ++c) avg[c] = sum[p][c] / (float)cnts[p];
        float sx[4] = {}, sy[4] = {}, sz[4] = {}, sw[4] = {};
#pragma unroll
        for (int wd = 0; wd < kExWords; ++wd)
            for (uint64_t m = mk[p][wd]; m; m &= m - 1) {
                const int t = wd * 64 + __builtin_ctzll(m);
                float px[4], d[4];
                PX(t, px);
                for (int c = 0; c < 4; ++c) d[c] = px[c] - avg[c];
                if (d[0] > 0) for (int c = 0; c < 4; ++c) sx[c] += d[c];
                if (d[1] > 0) for (int c = 0; c < 4; ++c) sy[c] += d[c];
                if (d[2] > 0) for (int c = 0; c < 4; ++c) sz[c] += d[c];
                if (rgba && d[3] > 0) for (int c = 0; c < 4; ++c) sw[c] += d[c];
            }
        float best[4];
        for (int c = 0; c < 4; ++c) best[c] = sx[c];
        if (rgba) {
            float bs = dot4(sx, sx), pr = dot4(sy, sy);
            if (pr > bs) { bs = pr; for (int c = 0; c < 4; ++c) best[c] = sy[c]; }
            pr = dot4(sz, sz);
            if (pr > bs) { bs = pr; for (int c = 0; c < 4; ++c) best[c] = sz[c]; }
            pr = dot4(sw, sw);
            if (pr > bs) { bs = pr; for (int c = 0; c < 4; ++c) best[c] = sw[c]; }
        } else {
            float bs = dot3(sx, sx), pr = dot3(sy, sy);
            if (pr > bs) { bs = pr; for (int c = 0; c < 4; ++c) best[c] = sy[c]; }
            pr = dot3(sz, sz);
            if (pr > bs) { bs = pr; for (int c = 0; c < 4; ++c) best[c] = sz[c]; }
            best[3] = 0.0f;
        }
        ex_normalize(avg, bavg[p]);
        ex_normalize(best, bdir[p]);
        const float ud = rgba ? dot4(avg, bdir[p]) : dot3(avg, bdir[p]);
        float amod[4];
        for (int c = 0; c < 4; ++c) amod[c] = avg[c] - bdir[p][c] * ud;
        float lo = 1e10f, hi = -1e10f;
#pragma unroll
        for (int wd = 0; wd < kExWords; ++wd)
            for (uint64_t m = mk[p][wd]; m; m &= m - 1) {
                const int t = wd * 64 + __builtin_ctzll(m);
                float px[4];
                PX(t, px);
                float up = (px[0] * bdir[p][0] + px[1] * bdir[p][1]) + px[2] * bdir[p][2];
                float sp = (px[0] * bavg[p][0] + px[1] * bavg[p][1]) + px[2] * bavg[p][2];
                if (rgba) {
                    up = up + px[3] * bdir[p][3];
                    sp = sp + px[3] * bavg[p][3];
                }
                lo = up < lo ? up : lo;
                hi = up > hi ? up : hi;
                float err = 0.0f, serr = 0.0f;
                for (int c = 0; c < (rgba ? 4 : 3); ++c) {
                    const float dd = (amod[c] - px[c]) + (up * bdir[p][c]);
                    const float ds = sp * bavg[p][c] - px[c];
                    err = c ? err + kCW[c] * dd * dd : kCW[c] * dd * dd;
                    serr = c ? serr + kCW[c] * ds * ds : kCW[c] * ds * ds;
                }
                ue += err;
                se += serr;
            }
        const float l = hi - lo;
        len[p] = l > 1e-7f ? l : 1e-7f;
    }
#pragma unroll
    for (int p = 0; p < PC; ++p) {
        const float ew = (float)cnts[p] * wie;
        float uv[4], sv[4];
        for (int c = 0; c < 4; ++c) { uv[c] = bdir[p][c] * len[p]; sv[c] = bavg[p][c] * len[p]; }
        ue += (rgba ? dot4(uv, uv) : dot3(uv, uv)) * ew;
        se += (rgba ? dot4(sv, sv) : dot3(sv, sv)) * ew;
    }
    *ue_out = ue;
    *se_out = se;
}

// Learned policy: the chunk blocks whose partition candidates count pc
// needs (plan_uses_positions, or with `seeded` plan_uses_seeds), compacted
// into list[0 .. *count) so that the exhaustive search's waves hold only
// such blocks.
// One counter update per 256 blocks (the waves' counts summed in LDS): the
// list's order is free, since every kernel that reads it handles each
// listed block on its own.
constexpr int kNeedListThreads = 256;
__global__ void __launch_bounds__(kNeedListThreads) k_need_list(ImageArgs A, SearchPlan P, int pc, int seeded, int* list,
                                                                int* count) {
    __shared__ int s_cnt[kNeedListThreads / 64];
    __shared__ int s_base;
    const int k = blockIdx.x * kNeedListThreads + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool need = k < A.nb && (seeded ? plan_uses_seeds(P, pc, k, A.blk0 + k) : plan_uses_positions(P, pc, k, A.blk0 + k));
    const unsigned long long m = __ballot(need);
    if (lane == 0) s_cnt[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = 0;
        for (int w = 0; w < kNeedListThreads / 64; ++w) total += s_cnt[w];
        s_base = total ? atomicAdd(count, total) : 0;
    }
    __syncthreads();
    int base = s_base;
    for (int w = 0; w < wave; ++w) base += s_cnt[w];
    if (need) list[base + __popcll(m & ((1ull << lane) - 1ull))] = k;
}

// Exhaustive lists longer than T.exw_max_list run on k_partition_ex (a lane
// per block), whose throughput per block is higher once its waves fill the
// chip; shorter ones on k_partition_exw (a wave per block).

// Texels staged as converted floats up to T.ex_f4_max_tc texels (4x4 by
// default: LDS per workgroup stays small enough for 6 workgroups per CU), as
// RGBA8 words above.
size_t partition_ex_lds_bytes(const DevTables& T) {
    const size_t per_texel = T.tc <= T.ex_f4_max_tc ? 16 : 4;
    return (size_t)T.tc * 64 * (per_texel + 1);
}

// kSeeded: the learned policy's variant.  The line fits cover the first
// P.mis_candidates partitionings of the mismatch order (the k-means and
// ordering are skipped when that is 0), then the block's P.seeds most likely
// seeds of this count from the predictor's seed head (in rank order; seeds
// that are not selected partitionings or were already fitted are skipped),
// and the 4 best are kept the same way.
template <int PC, bool kF4, bool kSeeded>
__global__ void __launch_bounds__(64) k_partition_ex(ImageArgs A, DevTables T, int16_t* pb_part, uint8_t* scratch,
                                                     const int* list, const int* list_count, int search_limit,
                                                     SearchPlan P) {
    constexpr int pc = PC;
    constexpr int kExWords = kF4 ? 1 : (kMaxTexels + 63) / 64;  // 64-texel words of a partition's texel mask
    extern __shared__ __align__(16) uint8_t smem[];
    const int tc = T.tc, lane = threadIdx.x, nb = A.nb;
    // scratch column of this lane (blocks of a list use their list slot;
    // short exhaustive lists are k_partition_exw's)
    const int col = blockIdx.x * 64 + lane;
    if (list && *list_count <= T.exw_max_list && (!kSeeded || P.mis_candidates == 0)) return;  // wave-per-block kernels
    const int k = list ? (col < *list_count ? list[col] : nb) : col;
    uint32_t* s_tex = reinterpret_cast<uint32_t*>(smem);             // [tc][64] RGBA8 (!kF4)
    float4* s_f4 = reinterpret_cast<float4*>(smem);                    // [tc][64] search values (kF4)
    uint8_t* s_part = smem + (size_t)tc * 64 * (kF4 ? 16 : 4);         // [tc][64] k-means cluster
    const int g = A.blk0 + k;
    if (k >= nb) return;  // no barriers below
    const BlockLoc loc = block_loc(A, T, g);
    uint32_t amin = 0xFFu, amax = 0u;
    for (int t = 0; t < tc; ++t) {
        const uint32_t v = texel_rgba8(loc, T, t);
        if (kF4)
            s_f4[t * 64 + lane] = make_float4(u8_search_value(v & 0xFF), u8_search_value((v >> 8) & 0xFF),
                                              u8_search_value((v >> 16) & 0xFF), u8_search_value(v >> 24));
        else
            s_tex[t * 64 + lane] = v;
        amin = min(amin, v >> 24);
        amax = max(amax, v >> 24);
    }
    auto PX = [&](int t, float* o) __attribute__((always_inline)) {
        if (kF4) {
            const float4 f = s_f4[t * 64 + lane];
            o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = f.w;
        } else {
            const uint32_t v = s_tex[t * 64 + lane];
            o[0] = u8_search_value(v & 0xFF);
            o[1] = u8_search_value((v >> 8) & 0xFF);
            o[2] = u8_search_value((v >> 16) & 0xFF);
            o[3] = u8_search_value(v >> 24);
        }
    };
    const int nsel = T.count_selected[pc - 1];
    uint16_t* order = nullptr;
    // seeded: the first mis_n partitionings of the mismatch order are
    // line-fitted before the seeds (0 = the seeds alone)
    const int mis_n = kSeeded ? min(P.mis_candidates, nsel) : 0;
    uint8_t* mis = scratch;
    if (!kSeeded || mis_n > 0) {
    // --- k-means: weighted-random seeding, then assign / update / assign /
    // update / assign (best_partitionings.cpp:26-194, 419-430).  The running
    // minimum distances are recomputed from the chosen centres (a min is exact)
    float centers[4][4];
    {
        float px[4];
        PX(145897 % tc, px);
        for (int c = 0; c < 4; ++c) centers[0][c] = px[c];
    }
#pragma unroll
    for (int n = 1; n < pc; ++n) {
        auto dmin = [&](int t) __attribute__((always_inline)) {
            float px[4];
            PX(t, px);
            float d = ex_dist(px, centers[0]);
#pragma unroll
            for (int j = 1; j < n; ++j) {
                const float dj = ex_dist(px, centers[j]);
                d = dj < d ? dj : d;
            }
            return d;
        };
        float dsum = 0.0f;
        for (int t = 0; t < tc; ++t) dsum += dmin(t);
        const float cut = dsum * kKmeansCutoffs[3 * (pc - 2) + n - 1];
        float summa = 0.0f;
        int next = 0;
        for (; next < tc; ++next) {
            summa += dmin(next);
            if (summa >= cut) break;
        }
        next = min(next, tc - 1);
        float px[4];
        PX(next, px);
        for (int c = 0; c < 4; ++c) centers[n][c] = px[c];
    }
#pragma unroll
    for (int n = pc; n < 4; ++n)
        for (int c = 0; c < 4; ++c) centers[n][c] = 0.0f;
    for (int it = 0; it < 3; ++it) {
        if (it > 0) {
            float sum[4][4] = {};
            int cnt[4] = {0, 0, 0, 0};
            for (int t = 0; t < tc; ++t) {
                float px[4];
                PX(t, px);
                const int p = s_part[t * 64 + lane];
#pragma unroll
                for (int q = 0; q < pc; ++q)
                    if (p == q) {
                        for (int c = 0; c < 4; ++c) sum[q][c] += px[c];
                        cnt[q]++;
                    }
            }
#pragma unroll
            for (int q = 0; q < pc; ++q) {
                const float sc = cnt[q] ? 1.0f / (float)cnt[q] : 0.0f;
                for (int c = 0; c < 4; ++c) centers[q][c] = cnt[q] ? sum[q][c] * sc : 0.0f;
            }
        }
        int cnt[4] = {0, 0, 0, 0};
        for (int t = 0; t < tc; ++t) {
            float px[4];
            PX(t, px);
            float bd = 3.402823466e38f;
            int bp = 0;
#pragma unroll
            for (int q = 0; q < pc; ++q) {
                const float d = ex_dist(px, centers[q]);
                if (d < bd) { bd = d; bp = q; }
            }
            s_part[t * 64 + lane] = (uint8_t)bp;
#pragma unroll
            for (int q = 0; q < pc; ++q) cnt[q] += bp == q;
        }
        bool problem;
        do {
            problem = false;
#pragma unroll
            for (int q = 0; q < pc; ++q)
                if (cnt[q] == 0) {
                    problem = true;
                    const int old = s_part[q * 64 + lane];
#pragma unroll
                    for (int r = 0; r < pc; ++r) cnt[r] -= r == old;
                    cnt[q]++;
                    s_part[q * 64 + lane] = (uint8_t)q;
                }
        } while (problem);
    }
    // --- mismatch of every selected partitioning, stable counting sort
    // (best_partitionings.cpp:325-397); the histogram's running sum stops
    // at the k-means texel count as in the reference, and ordering slots it
    // leaves unwritten read partitioning 0
    const int nk = min(tc, 64);
    uint64_t a[4] = {0, 0, 0, 0};
    for (int i = 0; i < nk; ++i) {
        const int p = s_part[T.kmeans_texels[i] * 64 + lane];
#pragma unroll
        for (int q = 0; q < pc; ++q) a[q] |= p == q ? (1ull << i) : 0ull;
    }
    // scratch columns: mismatch [nsel][nb] u8, ordering [nsel][nb] u16,
    // mismatch histogram [64][nb] u16
    const size_t o_order = ((size_t)nsel * nb + 255) & ~size_t(255);
    const size_t o_hist = o_order + (((size_t)nsel * nb * 2 + 255) & ~size_t(255));
    order = reinterpret_cast<uint16_t*>(scratch + o_order);
    uint16_t* hist = reinterpret_cast<uint16_t*>(scratch + o_hist);
#define s_hist_at(m) hist[(size_t)(m) * nb + col]
    for (int m = 0; m < 64; ++m) s_hist_at(m) = 0;
    for (int i = 0; i < nsel; ++i) {
        const uint64_t* bb = T.coverage + ((size_t)(pc - 2) * kPartitionings + i) * 4;
        auto Pp = [&](int x, int y) { return __popcll(a[x] ^ bb[y]); };
        int v;
        if (pc == 2) {
            v = min(Pp(0, 0) + Pp(1, 1), Pp(0, 1) + Pp(1, 0));
        } else if (pc == 3) {
            const int v0 = min(Pp(1, 1) + Pp(2, 2), Pp(1, 2) + Pp(2, 1)) + Pp(0, 0);
            const int v1 = min(Pp(1, 0) + Pp(2, 2), Pp(1, 2) + Pp(2, 0)) + Pp(0, 1);
            const int v2 = min(Pp(1, 0) + Pp(2, 1), Pp(1, 1) + Pp(2, 0)) + Pp(0, 2);
            v = min(v0, min(v1, v2));
        } else {
            const int mx23 = min(Pp(2, 2) + Pp(3, 3), Pp(2, 3) + Pp(3, 2));
            const int mx13 = min(Pp(2, 1) + Pp(3, 3), Pp(2, 3) + Pp(3, 1));
            const int mx12 = min(Pp(2, 1) + Pp(3, 2), Pp(2, 2) + Pp(3, 1));
            const int mx03 = min(Pp(2, 0) + Pp(3, 3), Pp(2, 3) + Pp(3, 0));
            const int mx02 = min(Pp(2, 0) + Pp(3, 2), Pp(2, 2) + Pp(3, 0));
            const int mx01 = min(Pp(2, 1) + Pp(3, 0), Pp(2, 0) + Pp(3, 1));
            const int v0 = Pp(0, 0) + min(Pp(1, 1) + mx23, min(Pp(1, 2) + mx13, Pp(1, 3) + mx12));
            const int v1 = Pp(0, 1) + min(Pp(1, 0) + mx23, min(Pp(1, 2) + mx03, Pp(1, 3) + mx02));
            const int v2 = Pp(0, 2) + min(Pp(1, 1) + mx03, min(Pp(1, 0) + mx13, Pp(1, 3) + mx01));
            const int v3 = Pp(0, 3) + min(Pp(1, 1) + mx02, min(Pp(1, 2) + mx01, Pp(1, 0) + mx12));
            v = min(v0, min(v1, min(v2, v3)));
        }
        const int m = v / 2;  // < 64 (half of at most 2 x 64 differing bits)
        mis[(size_t)i * nb + col] = (uint8_t)m;
        s_hist_at(m)++;
        order[(size_t)i * nb + col] = 0;
    }
    {
        uint16_t sum = 0;
        for (int m = 0; m < nk; ++m) {
            const uint16_t c = s_hist_at(m);
            s_hist_at(m) = sum;
            sum = (uint16_t)(sum + c);
        }
    }
    for (int i = 0; i < nsel; ++i) {
        const int m = mis[(size_t)i * nb + col];
        const int pos = s_hist_at(m);
        s_hist_at(m) = (uint16_t)(pos + 1);
        if (pos < nsel) order[(size_t)pos * nb + col] = (uint16_t)i;
        // seeded: mark the partitionings line-fitted from the order, so a
        // seed among them is not fitted twice
        if (kSeeded && pos < mis_n) mis[(size_t)i * nb + col] = (uint8_t)(m | 0x80);
    }
#undef s_hist_at
    }  // !kSeeded || mis_n > 0
    // --- line fits in mismatch order (find_best_partition_candidates,
    // compute_avgs_and_dirs_{4_comp,3_comp_rgb}, compute_error_squared_{rgba,rgb})
    const bool rgba = amin != amax;
    float wie = 0.055f;
    if (tc <= 20) wie = 0.03f;
    else if (tc <= 31) wie = 0.04f;
    else if (tc <= 41) wie = 0.05f;
    wie = wie * wie;
    float ube[4] = {1e30f, 1e30f, 1e30f, 1e30f}, sbe[4] = {1e30f, 1e30f, 1e30f, 1e30f};
    int ubi[4] = {0, 0, 0, 0}, sbi[4] = {0, 0, 0, 0};
    // partition_search_limit (best_partitionings.cpp:520): the line fits
    // cover the first search_limit partitionings of the mismatch order
    const int* seeds = kSeeded ? P.topk_idx + (size_t)k * 5 * P.kmax + (size_t)(pc - 1) * P.kmax : nullptr;
    const int nsearch = kSeeded ? mis_n + min(P.seeds, P.kmax) : min(nsel, search_limit);
    int nvalid = 0;
    for (int s = 0; s < nsearch; ++s) {
        int i;
        if (!kSeeded || s < mis_n) {
            i = order[(size_t)s * nb + col];
        } else {
            const int seed = seeds[s - mis_n];
            if (seed < 0 || seed >= kPartitionings) break;  // the head has no more columns
            i = T.seed_to_packed[(pc - 2) * kPartitionings + seed];
            if (i >= nsel) continue;  // not a selected partitioning of this count
            if (mis_n > 0 && (mis[(size_t)i * nb + col] & 0x80)) continue;  // fitted from the order
        }
        ++nvalid;
        float ue, se;
        ex_line_fit<PC, kExWords>(PX, T, pc, i, rgba, wie, &ue, &se);
        ex_insert(ue, i, ube, ubi);
        ex_insert(se, i, sbe, sbi);
    }
    const int req = min(nvalid, 4);
    int pos[4] = {-1, -1, -1, -1}, emitted = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int idx = (j & 1) ? sbi[j >> 1] : ubi[j >> 1];
        bool dup = emitted >= req || (j >> 1) >= req;
#pragma unroll
        for (int q = 0; q < 4; ++q) dup |= q < emitted && pos[q] == idx;
        if (!dup) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (q == emitted) pos[q] = idx;
            ++emitted;
        }
    }
    // a missing candidate (fewer than 4 distinct) repeats the first one: a
    // duplicate search never beats the first under the strict < of the
    // picks.  Seeded: -1 (k_plan lists only the candidates found)
    for (int c = 0; c < 4; ++c) pb_part[k * 4 + c] = (int16_t)(kSeeded || pos[c] >= 0 ? pos[c] : pos[0]);
}

// ---- pass01 + pass08 + pass09 + pass10 --------------------------------------
// One lane per partitioned block.  Per-partition accumulators are updated
// through a compile-time partition index (texel partition compared against
// each of the four), so they stay in registers; pass09 and pass10 run
// quant level by quant level.
__device__ __constant__ int kColorLevels[17] = {5, 7, 9, 11, 15, 19, 23, 31, 39, 47, 63, 79, 95, 127, 159, 191, 255};

// Specialised per partition count (launched once per count).
template <int PC>
__global__ void __launch_bounds__(64) k_ideal_color(ImageArgs A, DevTables T, const int16_t* pb_part, int npb,
                                                    Ideal* ideal_out, float* wts_out, float* ct_err, uint16_t* ct_fmt) {
    constexpr int pc = PC;
    const int pb = blockIdx.x * blockDim.x + threadIdx.x;
    if (pb >= pb_limit(A, pc, npb)) return;
    const Pb b = make_pb(A, T, pb_part, pc, pb);
    const int tc = T.tc;
    // partitioned blocks: the block's texel row (its 4 positions sit in
    // neighbouring lanes); one partition reads the image like its neighbours
    const uint32_t* row = A.blk_texels + (size_t)(b.g - A.blk0) * tc;
    auto px_at = [&](int i, float* o) {
        if (PC == 1) {
            texel_px(b.loc, T, i, o);
        } else {
            const uint32_t v = row[i];
            o[0] = u8_search_value(v & 0xFF);
            o[1] = u8_search_value((v >> 8) & 0xFF);
            o[2] = u8_search_value((v >> 16) & 0xFF);
            o[3] = u8_search_value(v >> 24);
        }
    };
    // pass01_ideal_endpoints_and_weights.wgsl:63-207
    const float error_weight = (((kCW[0] + kCW[1]) + kCW[2]) + kCW[3]) / 4.0f;
    float sum[4][4] = {}, count[4] = {};
    bool gray = true;
    float amin = 1e38f, amax = 0.0f;
    for (int i = 0; i < tc; ++i) {
        float px[4];
        px_at(i, px);
        const int p = texel_part(b, i);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (p == q) {
                for (int c = 0; c < 4; ++c) sum[q][c] += px[c];
                count[q] += 1.0f;
            }
        gray = gray && (px[0] == px[1] && px[0] == px[2]);
        amax = fmaxf(amax, px[3]);
        amin = fminf(amin, px[3]);
    }
    float avg[4][4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
        for (int c = 0; c < 4; ++c) avg[p][c] = count[p] > 0.0f ? sum[p][c] / count[p] : 0.0f;
    float dv[4][4] = {};
    for (int i = 0; i < tc; ++i) {
        float px[4];
        px_at(i, px);
        const int p = texel_part(b, i);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (p == q)
                for (int c = 0; c < 4; ++c) { const float d = wmax(px[c] - avg[q][c], 0.0f); dv[q][c] += d * d; }
    }
    float dir[4][4], minp[4], maxp[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) { normalize4(dv[p], dir[p]); minp[p] = 1e10f; maxp[p] = -1e10f; }
    for (int i = 0; i < tc; ++i) {
        float px[4];
        px_at(i, px);
        const int p = texel_part(b, i);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (p == q) {
                float df[4];
                for (int c = 0; c < 4; ++c) df[c] = px[c] - avg[q][c];
                const float proj = dot4(df, dir[q]);
                minp[q] = wmin(minp[q], proj);
                maxp[q] = wmax(maxp[q], proj);
            }
    }
    Ideal I;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        if (minp[p] >= maxp[p]) { minp[p] = 0.0f; maxp[p] = 1e-7f; }
        const float len = maxp[p] - minp[p];
        const float lensq = len * len;
        const bool zero = dir[p][0] == 0.0f && dir[p][1] == 0.0f && dir[p][2] == 0.0f && dir[p][3] == 0.0f;
        for (int c = 0; c < 4; ++c) {
            I.avg[p][c] = avg[p][c];
            I.ep0[p][c] = zero ? avg[p][c] : avg[p][c] + dir[p][c] * minp[p];
            I.ep1[p][c] = zero ? avg[p][c] : avg[p][c] + dir[p][c] * maxp[p];
        }
        I.wes[p] = lensq * error_weight;
    }
    for (int i = 0; i < tc; ++i) {
        float px[4];
        px_at(i, px);
        const int p = texel_part(b, i);
        float w = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (p == q) {
                float df[4];
                for (int c = 0; c < 4; ++c) df[c] = px[c] - avg[q][c];
                const float proj = dot4(df, dir[q]);
                const float span = wmax(maxp[q] - minp[q], 1e-6f);
                w = wclamp((proj - minp[q]) / span, 0.0f, 1.0f);
            }
        wts_out[(size_t)i * npb + pb] = w;  // [texel][PB]: coalesced
    }
    float min_ep[4] = {10.0f, 10.0f, 10.0f, 10.0f};
#pragma unroll
    for (int p = 0; p < 4; ++p)
        for (int c = 0; c < 4; ++c) {
            const float ep = (1.0f - I.ep0[p][c]) / ((I.ep1[p][c] - I.ep0[p][c]) + 1e-9f);
            if (ep > 0.5f && ep < min_ep[c]) min_ep[c] = ep;
        }
    I.cutoff = wmin(min_ep[0], wmin(min_ep[1], wmin(min_ep[2], min_ep[3])));
    I.pad[0] = I.pad[1] = I.pad[2] = 0.0f;
    ideal_out[pb] = I;

    // pass08_compute_encoding_choice_errors.wgsl:115-326
    float sxp[4][4] = {}, syp[4][4] = {}, szp[4][4] = {};
    for (int i = 0; i < tc; ++i) {
        float px[4];
        px_at(i, px);
        const int p = texel_part(b, i);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (p == q) {
                float t[4];
                for (int c = 0; c < 4; ++c) t[c] = px[c] - avg[q][c];
                t[3] = 0.0f;
                if (t[0] > 0.0f) for (int c = 0; c < 4; ++c) sxp[q][c] += t[c];
                if (t[1] > 0.0f) for (int c = 0; c < 4; ++c) syp[q][c] += t[c];
                if (t[2] > 0.0f) for (int c = 0; c < 4; ++c) szp[q][c] += t[c];
            }
    }
    float ub[4][4], ua[4][4], sb[4][4], la[4][4], lb[4];
    const float one[4] = {1.0f, 1.0f, 1.0f, 0.0f};
    normalize4(one, lb);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        for (int c = 0; c < 4; ++c) { ub[p][c] = 0.0f; ua[p][c] = 0.0f; sb[p][c] = 0.0f; la[p][c] = 0.0f; }
        if (p >= pc) continue;
        float best[4];
        float bs = dot4(sxp[p], sxp[p]);
        for (int c = 0; c < 4; ++c) best[c] = sxp[p][c];
        const float py = dot4(syp[p], syp[p]), pz = dot4(szp[p], szp[p]);
        if (py > bs) { for (int c = 0; c < 4; ++c) best[c] = syp[p][c]; bs = py; }
        if (pz > bs) { for (int c = 0; c < 4; ++c) best[c] = szp[p][c]; bs = pz; }
        const float a3[4] = {avg[p][0], avg[p][1], avg[p][2], 0.0f};
        normalize4(best, ub[p]);
        const float s = dot4(a3, ub[p]);
        for (int c = 0; c < 4; ++c) ua[p][c] = a3[c] - ub[p][c] * s;
        normalize4(a3, sb[p]);
        const float sl = dot4(a3, lb);
        for (int c = 0; c < 4; ++c) la[p][c] = a3[c] - lb[c] * sl;
    }
    float acc[4][5] = {};
    for (int i = 0; i < tc; ++i) {
        float px[4];
        px_at(i, px);
        const int p = texel_part(b, i);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (p == q) {
                const float ad = px[3] - 65536.0f;
                acc[q][0] += ad * ad * kCW[3];
                float d[3];
                float s = dot3(px, ub[q]);
                for (int c = 0; c < 3; ++c) { const float t = (ua[q][c] + s * ub[q][c]) - px[c]; d[c] = t * t; }
                acc[q][1] += dot3(d, kCW);
                s = dot3(px, sb[q]);
                for (int c = 0; c < 3; ++c) { const float t = s * sb[q][c] - px[c]; d[c] = t * t; }
                acc[q][2] += dot3(d, kCW);
                s = dot3(px, lb);
                for (int c = 0; c < 3; ++c) { const float t = (la[q][c] + s * lb[c]) - px[c]; d[c] = t * t; }
                acc[q][3] += dot3(d, kCW);
                for (int c = 0; c < 3; ++c) { const float t = s * lb[c] - px[c]; d[c] = t * t; }
                acc[q][4] += dot3(d, kCW);
            }
    }
    const bool constant_alpha = amax == amin;
    const int can_bc = !(gray && constant_alpha);

    // pass09_compute_color_error.wgsl:140-268: quant-independent terms
    const float lim = 0.12f * 65536.0f;
    float rse[4], le[4], ade[4], rgb_range[4], alpha_range[4], base_rgb[4], base_a[4], oe_rgba0[4], oe_rgb0[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        rse[p] = (acc[p][2] - acc[p][1]) * 0.7f;
        le[p] = (acc[p][4] - acc[p][1]) * 3.0f;
        ade[p] = acc[p][0] * 3.0f;
        bool can_oe = true;
        for (int c = 0; c < 3; ++c) can_oe = can_oe && (fabsf(I.ep1[p][c] - I.ep0[p][c]) < lim);
        const float size = (float)b.ppc[p];
        float sre[4];
        for (int c = 0; c < 4; ++c) {
            const float h0 = wmax(I.ep0[p][c] - 65535.0f, 0.0f), h1 = wmax(I.ep1[p][c] - 65535.0f, 0.0f);
            const float l0 = wmin(I.ep0[p][c], 0.0f), l1 = wmin(I.ep1[p][c], 0.0f);
            sre[c] = (((l0 * l0) + (l1 * l1)) + (h0 * h0)) + (h1 * h1);
        }
        rgb_range[p] = dot3(sre, kCW) * 0.5f * size;
        alpha_range[p] = sre[3] * kCW[3] * 0.5f * size;
        base_rgb[p] = ((kCW[0] + kCW[1]) + kCW[2]) * size;
        base_a[p] = kCW[3] * size;
        oe_rgba0[p] = can_oe ? 0.5f : 1.0f;
        oe_rgb0[p] = can_oe ? 0.25f : 1.0f;
    }
    const float bc_rgba = can_bc ? 0.625f : 1.0f, bc_rgb = can_bc ? 0.5f : 1.0f;
    // [quant - 4][column][PB]; rows for quant levels below 4 are never read
    auto ct_at = [&](int q, int col) { return ((size_t)(q - 4) * 13 + col) * npb + pb; };
    for (int i = 0; i < 17; ++i) {
        const int q = i + 4;
        const int K = kColorLevels[i];
        const float bqe = (65536.0f * 65536.0f / 18.0f) / (float)(K * K);
        float err[4][4];
        int fmt[4][4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const float oe_rgba = q >= 19 ? 1.0f : oe_rgba0[p], oe_rgb = q >= 19 ? 1.0f : oe_rgb0[p];
            const float qe_rgb = base_rgb[p] * bqe, qe_rgba = (base_rgb[p] + base_a[p]) * bqe;
            err[p][3] = ((qe_rgba * bc_rgba * oe_rgba) + rgb_range[p]) + alpha_range[p];
            fmt[p][3] = FMT_RGBA;
            const float full = ((qe_rgb * bc_rgb * oe_rgb) + rgb_range[p]) + ade[p];
            const float rgbs_a = ((qe_rgba + rse[p]) + rgb_range[p]) + alpha_range[p];
            if (rgbs_a < full) { err[p][2] = rgbs_a; fmt[p][2] = FMT_RGB_SCALE_ALPHA; }
            else { err[p][2] = full; fmt[p][2] = FMT_RGB; }
            const float rgbs = ((qe_rgb + rgb_range[p]) + ade[p]) + rse[p];
            const float lumal = ((qe_rgba + rgb_range[p]) + alpha_range[p]) + le[p];
            if (rgbs < lumal) { err[p][1] = rgbs; fmt[p][1] = FMT_RGB_SCALE; }
            else { err[p][1] = lumal; fmt[p][1] = FMT_LUMINANCE_ALPHA; }
            err[p][0] = ((qe_rgb + rgb_range[p]) + ade[p]) + le[p];
            fmt[p][0] = FMT_LUMINANCE;
        }
        if (pc == 1) {
            for (int k = 0; k < 4; ++k) { ct_err[ct_at(q, k)] = err[0][k]; ct_fmt[ct_at(q, k)] = (uint16_t)fmt[0][k]; }
            continue;
        }
        // pass10_color_combinations_for_quant_{2,3,4}part.wgsl
        float ce[13];
        uint16_t cf[13];
#pragma unroll
        for (int k = 0; k < 13; ++k) { ce[k] = kErrDefault; cf[k] = 0; }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int bb = 0; bb < 4; ++bb) {
                const int lo2 = min(a, bb), hi2 = max(a, bb);
                if (hi2 - lo2 > 1) continue;
                if (pc == 2) {
                    const float te = wmin(err[0][a] + err[1][bb], 1e10f);
                    if (te < ce[a + bb]) {
                        const int f[4] = {fmt[0][a], fmt[1][bb], 0, 0};
                        ce[a + bb] = te; cf[a + bb] = pack_fmt4(f);
                    }
                    continue;
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int lo3 = min(c, lo2), hi3 = max(c, hi2);
                    if (hi3 - lo3 > 1) continue;
                    if (pc == 3) {
                        const float te = wmin((err[0][a] + err[1][bb]) + err[2][c], 1e10f);
                        if (te < ce[a + bb + c]) {
                            const int f[4] = {fmt[0][a], fmt[1][bb], fmt[2][c], 0};
                            ce[a + bb + c] = te; cf[a + bb + c] = pack_fmt4(f);
                        }
                        continue;
                    }
#pragma unroll
                    for (int d = 0; d < 4; ++d) {
                        const int lo4 = min(d, lo3), hi4 = max(d, hi3);
                        if (hi4 - lo4 > 1) continue;
                        const float te = wmin(((err[0][a] + err[1][bb]) + err[2][c]) + err[3][d], 1e10f);
                        if (te < ce[a + bb + c + d]) {
                            const int f[4] = {fmt[0][a], fmt[1][bb], fmt[2][c], fmt[3][d]};
                            ce[a + bb + c + d] = te; cf[a + bb + c + d] = pack_fmt4(f);
                        }
                    }
                }
            }
        for (int k = 0; k < 13; ++k) { ct_err[ct_at(q, k)] = ce[k]; ct_fmt[ct_at(q, k)] = cf[k]; }
    }
}

// ---- passes 02-05, 06, 07, 11 -----------------------------------------------
// A 64-lane workgroup runs one decimation mode for 64 partitioned blocks, one
// block per lane, so loop bounds and table reads are uniform across the
// wave.  The blocks' ideal weights and per-texel error scales are staged in
// LDS, and so are each lane's decimated weights; all three are laid out
// [index][lane] (conflict-free).  Workgroups are dispatched round-robin over
// the 8 XCDs, so the grid is ordered to send every decimation mode of one
// block group to the same XCD, whose L2 then holds that group's data.
// Passes 03-05 run step by step: pass05 only needs its running minimum at
// the step counts of the quant levels the mode can use, so each level tracks
// its own (error, step, cut) triple, updated by the four rules in the order
// pass05 applies them; that is the same sequence of comparisons on the same
// values.
constexpr int kDmPbs = 64;
constexpr int kXcds = 8;

// LDS of one k_dm_bm workgroup: decimated weights and scratch per lane, plus
// the texel partition map for partitioned blocks.  The ideal texel weights
// are read from global memory (coalesced, cache-resident) instead, which
// lets more workgroups share a CU at the large footprints.
size_t dm_bm_lds_bytes(const DevTables& T, int wstride, int pc) {
    return 2 * 64 * (size_t)wstride * sizeof(float) + (pc > 1 ? (((size_t)kDmPbs * T.tc + 15) & ~size_t(15)) : 0);
}

unsigned dm_bm_grid(int npb, int ndm_launch) {
    const int groups = (npb + kDmPbs - 1) / kDmPbs;
    return (unsigned)(((groups + kXcds - 1) / kXcds) * kXcds * ndm_launch);
}

template <int PC>
__global__ void __launch_bounds__(64) k_dm_bm(ImageArgs A, DevTables T, const int16_t* pb_part, int npb,
                                              const Ideal* ideal, const float* wts_all, const float* ct_err,
                                              const uint16_t* ct_fmt, Ccr* ccr,
                                              uint8_t* qw_out, int wstride,
                                              const uint32_t* mode_mask, int mask_words, int dm_first,
                                              int dm_count) {
    constexpr int pc = PC;
    extern __shared__ __align__(16) uint8_t smem[];
    const int ndm = T.ndm, tc = T.tc, lane = threadIdx.x;
    const int wg = blockIdx.x;
    // workgroup -> (group of 64 PBs, decimation mode of this launch's group);
    // consecutive workgroups go to different XCDs, so one XCD's workgroups
    // share PB groups (and their rows in L2)
    const int slot = wg / kXcds, xcd = wg - slot * kXcds;
    const int grp = xcd + kXcds * (slot / dm_count);
    const int m = T.dm_order[dm_first + slot - (slot / dm_count) * dm_count];
    const int pb0 = grp * kDmPbs;
    const int npb_live = pb_limit(A, pc, npb);
    if (pb0 >= npb_live) return;
    float* s_dw = reinterpret_cast<float*>(smem);
    float* s_aux = s_dw + 64 * wstride;
    uint8_t* s_tpart = reinterpret_cast<uint8_t*>(s_aux + 64 * wstride);  // pc > 1 only
    if (pc > 1) {
        const int nl = min(kDmPbs, npb_live - pb0);
        for (int e = lane; e < kDmPbs * tc; e += 64) {
            const int t = e / kDmPbs, l = e - t * kDmPbs, pb = pb0 + l;
            if (l >= nl) continue;
            int kb, cb, slot;
            pb_decode(A, pc, pb, &kb, &cb, &slot);
            s_tpart[t * kDmPbs + l] = T.part_of_texel[((size_t)(pc - 2) * kPartitionings + pb_part[slot]) * tc + t];
        }
    }
    // the weight unquantisation rows (quant levels 0..11) for pass07's
    // per-lane lookups, from LDS instead of constant memory
    __shared__ uint8_t s_q2u[12 * 32];
    for (int e = lane; e < 12 * 32; e += 64) s_q2u[e] = c_qt.quant_to_unquant[e >> 5][e & 31];
    __syncthreads();
    const int pb = pb0 + lane;
    if (pb >= npb_live) return;
    // learned policy: only the block's allowed trial modes are evaluated (k_top
    // never reads the others); a decimation mode none of them uses is skipped
    const uint32_t* mask = nullptr;
    if (mode_mask) {
        int kb, cb, slot;
        pb_decode(A, pc, pb, &kb, &cb, &slot);
        mask = mode_mask + (size_t)kb * mask_words;
        bool any = false;
        for (int li = T.dm_bm_off[m]; li < T.dm_bm_off[m + 1]; ++li) {
            const int k = T.dm_bm_list[li];
            any |= ((mask[k >> 5] >> (k & 31)) & 1u) != 0u;
        }
        if (!any) return;
    }
    // ideal texel weights of this PB ([texel][PB]: coalesced across lanes)
    auto wts_of = [&](int t) { return wts_all[(size_t)t * npb + pb]; };
    // pass01 stores lengthSqr[p] * error_weight for every texel of partition p
    const float4 wes4 = *reinterpret_cast<const float4*>(ideal[pb].wes);
    const uint8_t* tpart = s_tpart + lane;
    auto wes_of = [&](int t) {
        const int p = pc > 1 ? tpart[t * 64] : 0;
        return p == 0 ? wes4.x : p == 1 ? wes4.y : p == 2 ? wes4.z : wes4.w;
    };
    float* dw = s_dw + lane;    // dw[i * 64]
    float* aux = s_aux + lane;  // aux[i * 64]
    const DmInfo di = T.dm[m];
    const int wc = di.weight_count;
    const uint8_t* t2wc = T.t2w_count + (size_t)m * tc;
    const uint8_t* t2wi = T.t2w_idx + (size_t)m * tc * 4;
    const float* t2ww = T.t2w_w + (size_t)m * tc * 4;
    const uint8_t* w2tc = T.w2t_count + m * 64;
    const uint16_t* w2to = T.w2t_off + m * 64;
    const uint8_t* w2ti = T.w2t_idx + di.w2t_base;
    const float* w2tw = T.w2t_w + di.w2t_base;
    int pb_in_batch;
    {
        int kb, c, slot;
        pb_decode(A, pc, pb, &kb, &c, &slot);
        const int j = block_loc(A, T, A.blk0 + kb).local % kRefBatch;
        pb_in_batch = pc == 1 ? j : j * 4 + c;
    }
    // pass02_decimated_weights.wgsl:132-270
    if (tc == wc) {
        for (int i = 0; i < wc; ++i) dw[i * 64] = wts_of(i);
    } else if (tc <= wstride) {
        // The lane's rows hold a whole block: cache the ideal weights in dw,
        // form the initial estimates in aux, then each texel's infill minus
        // its ideal weight once (the refinement step reads the initial
        // estimates only, so a texel's infill is the same for every weight
        // it touches), refine each weight in place in aux and copy the
        // results to dw.  Same float operations in the same order as the
        // loop below, without recomputing infills per weight-texel pair.
        for (int t = 0; t < tc; ++t) dw[t * 64] = wts_of(t);
        for (int w = 0; w < wc; ++w) {
            const int n = w2tc[w];
            float v = 0.0f;
            if (n != 0) {
                const int off = w2to[w];
                float ws = 0.0f, tcon = 1e-10f;
                for (int jx = 0; jx < n; ++jx) {
                    const int t = w2ti[off + jx];
                    const float cw = w2tw[off + jx] * wes_of(t);
                    ws += dw[t * 64] * cw;
                    tcon += cw;
                }
                v = ws / tcon;
            }
            aux[w * 64] = v;  // initial estimate
        }
        for (int t = 0; t < tc; ++t) {
            float infill = 0.0f;
            const int nt = t2wc[t];
            for (int k = 0; k < nt; ++k) infill += aux[t2wi[t * 4 + k] * 64] * t2ww[t * 4 + k];
            dw[t * 64] = infill - dw[t * 64];
        }
        for (int w = 0; w < wc; ++w) {
            float a = 1e-10f, bacc = 0.0f;
            const int n = w2tc[w], off = w2to[w];
            for (int jx = 0; jx < n; ++jx) {
                const int t = w2ti[off + jx];
                const float contrib = w2tw[off + jx];
                const float scale = wes_of(t) * contrib;
                a += contrib * scale;
                bacc += dw[t * 64] * scale;
            }
            aux[w * 64] = aux[w * 64] + wclamp((bacc * -16.0f) / a, -0.25f, 0.25f);
        }
        for (int w = 0; w < wc; ++w) dw[w * 64] = aux[w * 64];
    } else {
        // The block does not fit the lane's rows, so the per-weight sums are
        // taken in one walk over the texels, each texel adding its terms to
        // the weights it touches: a weight's texels are listed in increasing
        // texel order and both tables hold the same contribution, so every
        // sum takes the same terms in the same order as a walk over the
        // weight's own texels, with one ideal-weight load and one infill per
        // texel instead of one per weight-texel pair.
        // initial estimates: sums in aux, contribution totals in dw
        for (int w = 0; w < wc; ++w) { aux[w * 64] = 0.0f; dw[w * 64] = 1e-10f; }
        for (int t = 0; t < tc; ++t) {
            const int nt = t2wc[t];
            const float wt = wts_of(t), es = wes_of(t);
            for (int k = 0; k < nt; ++k) {
                const int w = t2wi[t * 4 + k];
                const float cw = t2ww[t * 4 + k] * es;
                aux[w * 64] += wt * cw;
                dw[w * 64] += cw;
            }
        }
        for (int w = 0; w < wc; ++w) aux[w * 64] = w2tc[w] != 0 ? aux[w * 64] / dw[w * 64] : 0.0f;
        // one refinement step per weight against the infill of the initial
        // estimate: the gradient sums in dw, then each weight's curvature
        for (int w = 0; w < wc; ++w) dw[w * 64] = 0.0f;
        for (int t = 0; t < tc; ++t) {
            const int nt = t2wc[t];
            float infill = 0.0f;
            for (int k = 0; k < nt; ++k) infill += aux[t2wi[t * 4 + k] * 64] * t2ww[t * 4 + k];
            const float r = infill - wts_of(t), es = wes_of(t);
            for (int k = 0; k < nt; ++k) dw[t2wi[t * 4 + k] * 64] += r * (es * t2ww[t * 4 + k]);
        }
        for (int w = 0; w < wc; ++w) {
            float a = 1e-10f;
            const int n = w2tc[w], off = w2to[w];
            for (int jx = 0; jx < n; ++jx) {
                const float contrib = w2tw[off + jx];
                a += contrib * (wes_of(w2ti[off + jx]) * contrib);
            }
            dw[w * 64] = aux[w * 64] + wclamp((dw[w * 64] * -16.0f) / a, -0.25f, 0.25f);
        }
    }
    // passes 03-05 (pass03_compute_angular_offsets.wgsl:65-116,
    // pass04_lowest_and_highest_weight.wgsl:116-227,
    // pass05_best_values_for_quant_levels.wgsl:84-175)
    float minw = 3.4e38f, maxw = -3.4e38f;
    for (int i = 0; i < wc; ++i) {
        const float v = dw[i * 64];
        if (v < minw) minw = v;
        if (v > maxw) maxw = v;
    }
    const int nq = di.max_quant_level + 1;  // <= 8 (QUANT_12)
    float be[8], bstep[8], brc[8], blow[8], bao[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) { be[q] = 1e37f; bstep[q] = 0.0f; brc[q] = 0.0f; blow[q] = 0.0f; bao[q] = 0.0f; }
    const float inv2pi = 1.0f / (2.0f * 3.14159265358979323846f);
    const int mqs = di.max_quant_steps;
    // pass03 for all 16 step counts at once: one 64-byte table row per weight
    // and table, each step's sums still taken in weight order
    float ao[16];
    {
        float ax[16], ay[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) { ax[s] = 0.0f; ay[s] = 0.0f; }
        for (int jx = 0; jx < wc; ++jx) {
            const uint32_t is = wf2u(wround(wclamp(dw[jx * 64], 0.0f, 1.0f) * (1024.0f - 1.0f)));
            const float4* cr = reinterpret_cast<const float4*>(T.cos_t + is * 16);
            const float4* sr = reinterpret_cast<const float4*>(T.sin_t + is * 16);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float4 c = cr[v], sn = sr[v];
                ax[4 * v + 0] += c.x; ax[4 * v + 1] += c.y; ax[4 * v + 2] += c.z; ax[4 * v + 3] += c.w;
                ay[4 * v + 0] += sn.x; ay[4 * v + 1] += sn.y; ay[4 * v + 2] += sn.z; ay[4 * v + 3] += sn.w;
            }
        }
#pragma unroll
        for (int s = 0; s < 16; ++s) ao[s] = atan2f_fixed(ay[s], ax[s]) * inv2pi;
    }
    for (int sp = 0; sp < di.max_angular_steps; ++sp) {
        float off = ao[0];
#pragma unroll
        for (int s = 1; s < 16; ++s)
            if (sp == s) off = ao[s];
        const float rcp = (float)(sp + 1);
        const float minidx = wround(minw * rcp - off), maxidx = wround(maxw * rcp - off);
        float e = 0.0f, cl = 0.0f, ch = 0.0f;
        for (int jx = 0; jx < wc; ++jx) {
            const float sval = dw[jx * 64] * rcp - off;
            const float r = wround(sval);
            const float diff = sval - r;
            e += diff * diff;
            cl += (r == minidx) ? 1.0f - (2.0f * diff) : 0.0f;
            ch += (r == maxidx) ? 1.0f + (2.0f * diff) : 0.0f;
        }
        int span = wf2i(maxidx - minidx + 1.0f);
        span = max(2, min(span, mqs + 3));
        const float ssize = 1.0f / rcp, es = ssize * ssize;
        const float he = e * es, hcl = cl * es, hch = ch * es;
        const float ecl = he + hcl, ech = he + hch, eclh = (he + hcl) + hch;
        const float fi = (float)sp;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (q >= nq) continue;
            if (sp == 0) { blow[q] = minidx; bao[q] = off; }  // step 0 stands in for an unset slot
            const int x = kSteps[q];
            bool upd = false;
            if (span == x && he < be[q]) { be[q] = he; bstep[q] = fi; brc[q] = 0.0f; upd = true; }
            if (span - 1 == x && ecl < be[q]) { be[q] = ecl; bstep[q] = fi; brc[q] = 1.0f; upd = true; }
            if (span - 1 == x && ech < be[q]) { be[q] = ech; bstep[q] = fi; brc[q] = 0.0f; upd = true; }
            if (span - 2 == x && eclh < be[q]) { be[q] = eclh; bstep[q] = fi; brc[q] = 1.0f; upd = true; }
            if (upd) { blow[q] = minidx; bao[q] = off; }
        }
    }
    // the 13-float output stride against an 8-float allocation
    // (shaders_init.cpp:1176): slots beyond 8*920*4*ndm floats read as 0
    float low[8], high[8];
    {
        const long long alloc = 8ll * kRefBatch * 4 * ndm;
        const long long t = (long long)pb_in_batch * ndm + m;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            low[q] = 0.0f;
            high[q] = 0.0f;
            if (q >= nq) continue;
            const int qs = kSteps[q];
            const int bsi = max(wf2i(bstep[q]), 0);
            const float lwi = blow[q] + brc[q];
            const float hwi = (lwi + (float)qs) - 1.0f;
            const float stepsize = 1.0f / (1.0f + (float)bsi);
            if (13 * t + q < alloc) {
                low[q] = (bao[q] + lwi) * stepsize;
                high[q] = (bao[q] + hwi) * stepsize;
            }
        }
    }
    // passes 06, 07 and 11 for every trial block mode of this decimation mode
    const Ideal& I = ideal[pb];
    const float cutoff = I.cutoff;
    // colour-format table entries of this PB: [quant - 4][column][PB]
    auto ct_at = [&](int q, int col) { return ((size_t)(q - 4) * 13 + col) * npb + pb; };
    for (int li = T.dm_bm_off[m]; li < T.dm_bm_off[m + 1]; ++li) {
        const int k = T.dm_bm_list[li];
        if (mask && !((mask[k >> 5] >> (k & 31)) & 1u)) continue;
        const BmInfo bmi = T.bm[k];
        Ccr out;
        out.ql = -1; out.ql_mod = -1; out.formats = 0;
        // the mode's region, this PB's row (the weight count rounded up to 4)
        const int wcp = (di.weight_count + 3) & ~3;
        uint32_t* qwo = reinterpret_cast<uint32_t*>(qw_out + (size_t)T.qw_off[k] * npb + (size_t)pb * wcp);
        const int q = bmi.quant;
        const int bc = kFreeBits[pc - 1] - bmi.weight_bits;
        float we;
        if (q > 11 || bc <= 0) {
            we = 1e38f;
        } else {
            // pass06_remap_low_and_high_values.wgsl:72-98 (quant <= 12 here)
            float lo = 0.0f, hi = 0.0f;
#pragma unroll
            for (int qq = 0; qq < 8; ++qq)
                if (qq == q) { lo = low[qq]; hi = high[qq]; }
            // pass07_weights_and_error_for_bm.wgsl:228-331
            if (hi > 1.02f * cutoff) hi = 1.0f;
            if (hi <= lo) { lo = 0.0f; hi = 1.0f; }
            const float qlm1 = (float)(kLevelsD[q] - 1);
            float rscale = hi - lo;
            const float scale = 1.0f / rscale;
            const float slb = lo * scale;
            rscale = rscale / 64.0f;
            const uint32_t qmax = wf2u(qlm1);
            // only whole groups of four weights reach the output (pass07:322-331)
            const int keep = (wc / 4) * 4;
            uint32_t word = 0;
            // an undecimated grid's weight i is texel i's whole infill (one
            // contribution of 1.0), and dw still holds the ideal weights, so
            // texel i's error is summed here, in the same order, with the
            // same value: (0 + q * 1) - w differs from q - w at most in the
            // sign of a zero, which squaring removes
            const bool undecimated = tc == wc;
            float err = 0.0f;
            for (int i = 0; i < wc; ++i) {
                const float ix = wclamp(dw[i * 64] * scale - slb, 0.0f, 1.0f);
                const float ix1 = ix * qlm1;
                const uint32_t wl = wf2u(ix1);
                const uint32_t wh = min(wl + 1u, qmax);
                const int il = s_q2u[q * 32 + wl], ih = s_q2u[q * 32 + wh];
                const float fl = (float)il, fh = (float)ih;
                const bool use_high = (fl + fh) < (128.0f * ix);
                word |= (uint32_t)(i < keep ? (use_high ? ih : il) : 0) << (8 * (i & 3));
                if ((i & 3) == 3 || i == wc - 1) {
                    qwo[i >> 2] = word;
                    word = 0;
                }
                const float qv = (use_high ? fh : fl) * rscale + lo;
                if (undecimated) {
                    const float diff = qv - dw[i * 64];
                    err += diff * diff * wes_of(i);
                } else {
                    aux[i * 64] = qv;
                }
            }
            for (int i = 0; i < tc && !undecimated; ++i) {
                float v = 0.0f;
                for (int jx = 0; jx < t2wc[i]; ++jx) v += aux[t2wi[i * 4 + jx] * 64] * t2ww[i * 4 + jx];
                const float diff = v - wts_of(i);
                err += diff * diff * wes_of(i);
            }
            we = err;
        }
        // pass11_best_color_combination_for_mode_{1,2,3,4}part.wgsl
        if (!(we < kErrDefault)) {
            out.total_error = kErrDefault;
        } else if (pc == 1) {
            int best_idx = 0;
            float best = kErrDefault;
            for (int ic = 1; ic <= 4; ++ic) {
                const int ql = c_qt.quant_mode[ic][bc];
                if (ql < 4) continue;
                const float e = ct_err[ct_at(ql, ic - 1)];
                if (e < best) { best = e; best_idx = ic; }
            }
            const int fql = c_qt.quant_mode[best_idx][bc];
            const int f = fql >= 4 ? ct_fmt[ct_at(fql, best_idx - 1)] : FMT_LUMINANCE;
            out.total_error = best + we;
            out.ql = (int8_t)fql;
            out.ql_mod = (int8_t)fql;
            out.formats = (uint16_t)f;
        } else {
            const int start = pc, end = pc == 2 ? 8 : 9, modbits = pc == 2 ? 2 : (pc == 3 ? 5 : 8);
            int best_idx = 0;
            float best = kErrDefault;
            for (int ic = start; ic <= end; ++ic) {
                const int ql = c_qt.quant_mode[ic][bc];
                if (ql < 4) break;
                const float e = ct_err[ct_at(ql, ic - start)];
                if (e < best) { best = e; best_idx = ic; }
            }
            const int fql = c_qt.quant_mode[best_idx][bc];
            const int fqlm = c_qt.quant_mode[best_idx][bc + modbits];
            out.formats = fql >= 4 ? ct_fmt[ct_at(fql, best_idx - start)] : 0;
            out.total_error = best + we;
            out.ql = (int8_t)fql;
            out.ql_mod = (int8_t)fqlm;
        }
        ccr[(size_t)k * npb + pb] = out;  // [mode][block]: coalesced across lanes
    }
}

// ---- pass12 -------------------------------------------------------------------
// Learned policy, refinement pruning: each partitioned block's lowest
// estimated candidate error over its allowed trial modes.
__global__ void __launch_bounds__(64) k_pb_est(ImageArgs A, int pc, int npb, int nbm, const Ccr* ccr,
                                               const uint32_t* mode_mask, int mask_words, float* est) {
    const int pb = blockIdx.x * blockDim.x + threadIdx.x;
    if (pb >= pb_limit(A, pc, npb)) return;
    int kb, cb, sl;
    pb_decode(A, pc, pb, &kb, &cb, &sl);
    const uint32_t* mask = mode_mask ? mode_mask + (size_t)kb * mask_words : nullptr;
    float best = kErrDefault;
    for (int k = 0; k < nbm; ++k) {
        if (mask && !((mask[k >> 5] >> (k & 31)) & 1u)) continue;
        best = fminf(best, ccr[(size_t)k * npb + pb].total_error);
    }
    est[pb] = best;
}

// Rank of partitioned block pb among the searched positions of its block by
// estimated error (lower first, ties to the lower position).
__device__ __forceinline__ int pb_est_rank(const ImageArgs& A, int pc, int pb, const float* est) {
    int kb, cb, sl;
    pb_decode(A, pc, pb, &kb, &cb, &sl);
    const float mine = est[pb];
    int r = 0;
    for (int c = 0; c < 4; ++c) {
        if (c == cb) continue;
        const int q = A.pb_inv ? A.pb_inv[kb * 4 + c] : kb * 4 + c;
        if (q < 0) continue;
        const float e = est[q];
        r += (e < mine) || (e == mine && c < cb);
    }
    return r;
}

// Top-8 trial modes per partitioned block, plus the per-mode candidate counts
// that size the refinement work list (LDS histogram, one global add per
// workgroup and mode).  Slots without a candidate get the default error.
// k_top and k_scatter: partitioned blocks per workgroup (a workgroup clears
// and flushes one LDS histogram over all buckets, so larger workgroups do
// that once for more blocks)
constexpr int kTopThreads = 256;

// Work-list class of a candidate: pass14 packs each partition's endpoints
// along the path of its requested format (luminance, luminance + alpha, RGB
// scale, RGB scale + alpha, RGB, RGBA); the RGB paths try the delta and
// blue-contraction encodings only below colour quant levels 19 / 20, and a
// block whose partitions request different formats walks several paths.
__device__ __forceinline__ int refine_class(const Ccr& c, int pc) {
    const int f0 = c.formats & 0xF;
    bool uniform = true;
    for (int p = 1; p < pc; ++p) uniform = uniform && fmt_of(c.formats, p) == f0;
    int fam;
    switch (f0) {
        case FMT_LUMINANCE_ALPHA: case FMT_LUMINANCE_ALPHA_DELTA: fam = 1; break;
        case FMT_RGB_SCALE: fam = 2; break;
        case FMT_RGB_SCALE_ALPHA: fam = 3; break;
        case FMT_RGB: case FMT_RGB_DELTA: fam = 4; break;
        case FMT_RGBA: case FMT_RGBA_DELTA: fam = 5; break;
        default: fam = 0; break;
    }
    if (!uniform) return 10 + fam;  // 10..15
    if (fam < 4) return fam;        // 0..3
    const int qc = c.ql <= 18 ? 0 : (c.ql == 19 ? 1 : 2);
    return 4 + 3 * (fam - 4) + qc;  // 4..9
}

__global__ void __launch_bounds__(kTopThreads) k_top(ImageArgs A, int pc, int npb, int nbm, const Ccr* ccr, int16_t* top,
                                            int* bin_count, CandOut* cand, float* cand_err, const uint32_t* mode_mask,
                                            int mask_words, const uint8_t* slots, const float* pb_est, int keep_refine,
                                            const uint8_t* fallback, int classes) {
    extern __shared__ int s_hist[];  // [nbm * kRefineClasses]
    const int nbins = nbm * kRefineClasses;
    for (int i = threadIdx.x; i < nbins; i += kTopThreads) s_hist[i] = 0;
    __syncthreads();
    const int pb = blockIdx.x * blockDim.x + threadIdx.x;
    if (pb < pb_limit(A, pc, npb)) {
        // learned policy: the block's allowed trial modes and slot count
        const uint32_t* mask = nullptr;
        int nslots = 8;
        if (mode_mask || slots) {
            int kb, cb, sl;
            pb_decode(A, pc, pb, &kb, &cb, &sl);
            if (mode_mask) mask = mode_mask + (size_t)kb * mask_words;
            if (slots) nslots = slots[kb];
        }
        if (pb_est) {  // refinement pruning, except for blocks that fall back to the full search
            int kb, cb, sl;
            pb_decode(A, pc, pb, &kb, &cb, &sl);
            if (!fallback[kb] && pb_est_rank(A, pc, pb, pb_est) >= keep_refine) nslots = 0;
        }
        float te[8];
        int ti[8];
        for (int i = 0; i < 8; ++i) { te[i] = kErrDefault; ti[i] = 0; }
        for (int k = 0; k < nbm; ++k) {
            if (mask && !((mask[k >> 5] >> (k & 31)) & 1u)) continue;
            const float e = ccr[(size_t)k * npb + pb].total_error;
            if (e < te[7]) {
                te[7] = e; ti[7] = k;
                for (int j = 7; j > 0; --j)
                    if (te[j] < te[j - 1]) {
                        const float fe = te[j]; te[j] = te[j - 1]; te[j - 1] = fe;
                        const int fi = ti[j]; ti[j] = ti[j - 1]; ti[j - 1] = fi;
                    }
            }
        }
        for (int i = 0; i < 8; ++i) {
            int v = (te[i] < kErrDefault && i < nslots) ? ti[i] : -1;
            if (v >= 0) v = (v << kRefineClassBits) | (classes ? refine_class(ccr[(size_t)v * npb + pb], pc) : 0);
            top[pb * 8 + i] = (int16_t)v;
            if (v >= 0) atomicAdd(&s_hist[v], 1);
            else {  // pass12 produced no candidate for this slot
                cand[pb * 8 + i].error = kErrDefault;
                cand_err[pb * 8 + i] = kErrDefault;
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += kTopThreads)
        if (s_hist[i]) atomicAdd(&bin_count[i], s_hist[i]);
}

// Work-list layout: each bucket (trial mode, class) starts at a multiple of
// 64; padding entries are -1.  Each lane lays out a contiguous range of
// buckets.  Resets the counts for the next launch.
__global__ void __launch_bounds__(64) k_bins(int nbins, int* bin_count, int* bin_cur, int* work, int* work_total,
                                             int* tile_next) {
    __shared__ int s_off[64];
    const int per = (nbins + 63) / 64;
    const int lo = min(nbins, (int)threadIdx.x * per), hi = min(nbins, lo + per);
    int sum = 0;
    for (int b = lo; b < hi; ++b) sum += (bin_count[b] + 63) & ~63;
    s_off[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int i = 0; i < 64; ++i) {
            const int v = s_off[i];
            s_off[i] = run;
            run += v;
        }
        *work_total = run;
        *tile_next = 0;
    }
    __syncthreads();
    int run = s_off[threadIdx.x];
    for (int b = lo; b < hi; ++b) {
        const int c = bin_count[b], padded = (c + 63) & ~63;
        bin_cur[b] = run;
        for (int i = run + c; i < run + padded; ++i) work[i] = -1;
        bin_count[b] = 0;
        run += padded;
    }
}

// Fills the buckets: positions are reserved per workgroup and bucket, so the
// order inside a bucket is arbitrary (every task writes its own output
// slot) but keeps each workgroup's partitioned blocks together.  Where the
// refinement stages texels, 256 partitioned blocks per workgroup (a tile of
// the refinement then reads the candidate state of fewer, neighbouring
// blocks: 4x4 COMPAT refinement 171 -> 164 ms); unstaged, 64 (its texel
// reads lose locality with the wider ranges: 6x6 learned 67 -> 71 ms).
__global__ void __launch_bounds__(kTopThreads) k_scatter(ImageArgs A, int pc, int npb, int nbins, const int16_t* top,
                                                int* bin_cur, int* work) {
    extern __shared__ int s_sc[];  // [2][nbins]
    int* s_cnt = s_sc;
    int* s_base = s_sc + nbins;
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) s_cnt[i] = 0;
    __syncthreads();
    const int pb = blockIdx.x * blockDim.x + threadIdx.x;
    int kk[8], loc[8];
    for (int s = 0; s < 8; ++s) {
        kk[s] = pb < pb_limit(A, pc, npb) ? (int)top[pb * 8 + s] : -1;
        loc[s] = kk[s] >= 0 ? atomicAdd(&s_cnt[kk[s]], 1) : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
        const int c = s_cnt[i];
        s_base[i] = c ? atomicAdd(&bin_cur[i], c) : 0;
    }
    __syncthreads();
    for (int s = 0; s < 8; ++s)
        if (kk[s] >= 0) work[s_base[kk[s]] + loc[s]] = pb * 8 + s;
}

// ---- pass14 helpers (pass14_pack_color_endpoints.wgsl:391-1221) -------------
__device__ __forceinline__ int u2q(int q, int index) { return c_qt.color_u2q[q - 4][index]; }

__device__ void quant3_int(int q, const int* v, int* o) {
    for (int c = 0; c < 3; ++c) o[c] = u2q(q, v[c] * 2 + 1);
    o[3] = 0;
}
__device__ void quant3_flt(int q, const int* v, const float* vf, int* o) {
    for (int c = 0; c < 3; ++c) {
        int idx = v[c] * 2;
        if (vf[c] - (float)v[c] >= -0.1f) idx += 1;
        o[c] = u2q(q, idx);
    }
    o[3] = 0;
}
__device__ __forceinline__ int quant1_int(int q, int v) { return u2q(q, v * 2 + 1); }
__device__ __forceinline__ int quant1_flt(int q, int v, float vf) {
    int idx = v * 2;
    if (vf - (float)v >= -0.1f) idx += 1;
    return u2q(q, idx);
}
__device__ void bit_transfer_signed(int* in0, int* in1) {
    for (int c = 0; c < 4; ++c) {
        const uint32_t a = (uint32_t)in0[c], b = (uint32_t)in1[c];
        const int nb = (int)((b >> 1) | (a & 0x80u));
        int na = (int)((a >> 1) & 0x3Fu);
        if (na & 0x20) na -= 0x40;
        in0[c] = na;
        in1[c] = nb;
    }
}
__device__ __forceinline__ void uncontract(int* v) {
    const int b = v[2];
    const int r = (v[0] + b) >> 1, g = (v[1] + b) >> 1;
    v[0] = r; v[1] = g;
}
__device__ void rgba_delta_unpack(const int* in0, const int* in1, int* o0, int* o1) {
    int i0[4], i1[4];
    for (int c = 0; c < 4; ++c) { i0[c] = in0[c]; i1[c] = in1[c]; }
    bit_transfer_signed(i1, i0);
    const int rgb_sum = i1[0] + i1[1] + i1[2];
    for (int c = 0; c < 4; ++c) i1[c] = i1[c] + i0[c];
    if (rgb_sum < 0) {
        uncontract(i0); uncontract(i1);
        for (int c = 0; c < 4; ++c) { const int t = i0[c]; i0[c] = i1[c]; i1[c] = t; }
    }
    for (int c = 0; c < 4; ++c) { o0[c] = min(max(i0[c], 0), 255); o1[c] = min(max(i1[c], 0), 255); }
}
__device__ void rgba_unpack(const int* in0, const int* in1, int* o0, int* o1) {
    int i0[4], i1[4];
    for (int c = 0; c < 4; ++c) { i0[c] = in0[c]; i1[c] = in1[c]; }
    if ((i0[0] + i0[1] + i0[2]) > (i1[0] + i1[1] + i1[2])) {
        uncontract(i0); uncontract(i1);
        for (int c = 0; c < 4; ++c) { const int t = i0[c]; i0[c] = i1[c]; i1[c] = t; }
    }
    for (int c = 0; c < 4; ++c) { o0[c] = i0[c]; o1[c] = i1[c]; }
}
__device__ float enc_error(const float* u0, const float* u1, const int* q0, const int* q1) {
    float e[4];
    for (int c = 0; c < 4; ++c) {
        const float e0 = u0[c] - (float)q0[c], e1 = u1[c] - (float)q1[c];
        e[c] = e0 * e0 + e1 * e1;
    }
    return ((e[0] + e[1]) + e[2]) + e[3];
}
__device__ __forceinline__ bool out_of_range(const float* c) {
    return c[0] < 0.0f || c[0] > 255.0f || c[1] < 0.0f || c[1] > 255.0f || c[2] < 0.0f || c[2] > 255.0f ||
           c[3] < 0.0f || c[3] > 255.0f;
}
__device__ bool try_rgb_delta_core(const float* c0, const float* c1, int* out0, int* out1, int q, bool blue) {
    int c0a[4], c0b[4], c0be[4], c1d[4], c1de[4];
    for (int i = 0; i < 4; ++i) c0a[i] = wf2i(wround(c0[i])) << 1;
    for (int i = 0; i < 4; ++i) c0b[i] = c0a[i] & 0xFF;
    quant3_int(q, c0b, c0be);
    for (int i = 0; i < 4; ++i) c0b[i] = c0be[i] | (c0a[i] & 0x100);
    for (int i = 0; i < 4; ++i) c1d[i] = (wf2i(wround(c1[i])) << 1) - c0b[i];
    c1d[3] = 0;
    for (int i = 0; i < 4; ++i) if (c1d[i] > 63 || c1d[i] < -64) return false;
    for (int i = 0; i < 4; ++i) c1d[i] = (c1d[i] & 0x7F) | ((c0b[i] & 0x100) >> 1);
    quant3_int(q, c1d, c1de);
    for (int i = 0; i < 3; ++i) if (((c1d[i] ^ c1de[i]) & 0xC0) != 0) return false;
    int ep0[4], ep1[4];
    for (int i = 0; i < 4; ++i) { ep0[i] = c0be[i]; ep1[i] = c1de[i]; }
    bit_transfer_signed(ep1, ep0);
    const int s = ep1[0] + ep1[1] + ep1[2];
    if (blue ? s >= 0 : s < 0) return false;
    for (int i = 0; i < 4; ++i) {
        ep0[i] = ep0[i] + ep1[i];
        if (ep0[i] < 0 || ep0[i] > 0xFF) return false;
    }
    for (int i = 0; i < 4; ++i) { out0[i] = c0be[i]; out1[i] = c1de[i]; }
    return true;
}
// try_quantize_rgb_delta_blue_contract, pass14:525-606
__device__ bool try_rgb_delta_bc(const float* color0, const float* color1, int* out0, int* out1, int q) {
    float c0[4], c1[4];
    const float s0[4] = {color1[2], color1[2], color1[2], color1[3]};
    const float s1[4] = {color0[2], color0[2], color0[2], color0[3]};
    for (int i = 0; i < 4; ++i) { c0[i] = (color1[i] + color1[i]) - s0[i]; c1[i] = (color0[i] + color0[i]) - s1[i]; }
    if (out_of_range(c0) || out_of_range(c1)) return false;
    return try_rgb_delta_core(c0, c1, out0, out1, q, true);
}
// try_quantize_rgb_blue_contract, pass14:674-711
__device__ bool try_rgb_bc(const float* color0, const float* color1, int* out0, int* out1, int q) {
    float c0[4], c1[4];
    const float s0[4] = {color0[2], color0[2], color0[2], color0[3]};
    const float s1[4] = {color1[2], color1[2], color1[2], color1[3]};
    for (int i = 0; i < 4; ++i) { c0[i] = (color0[i] + color0[i]) - s0[i]; c1[i] = (color1[i] + color1[i]) - s1[i]; }
    if (out_of_range(c0) || out_of_range(c1)) return false;
    int r0[4], r1[4], c0i[4], c1i[4];
    for (int i = 0; i < 4; ++i) { r0[i] = wf2i(wround(c0[i])); r1[i] = wf2i(wround(c1[i])); }
    quant3_flt(q, r0, c0, c0i);
    quant3_flt(q, r1, c1, c1i);
    if ((c1i[0] + c1i[1] + c1i[2]) <= (c0i[0] + c0i[1] + c0i[2])) return false;
    for (int i = 0; i < 4; ++i) { out0[i] = c1i[i]; out1[i] = c0i[i]; }
    return true;
}
// quantize_rgb, pass14:715-746.  The reference steps the endpoints apart,
// c0 down and c1 up by 0.2 per step (float subtractions / additions), until
// the quantised colour sums are ordered, and returns the quantised endpoints
// of that step.  For inverted endpoints that takes hundreds of steps, each
// with six table lookups.  Both sums are monotone in the step count (every
// float step moves a component strictly, the rounded index 2*round(c) +
// [c - round(c) >= -0.1] follows it, and the table is non-decreasing in its
// index), so the first ordered step is found by doubling the step count and
// then bisecting; the endpoints at a step are recomputed by repeating the
// same float operations from the last unordered step, so the values, and the
// result, are those of the sequential loop.
__device__ __forceinline__ void rgb_steps(float* c0, float* c1, int n) {
    for (int s = 0; s < n; ++s)
        for (int i = 0; i < 3; ++i) {
            c0[i] = c0[i] - 0.2f;
            c1[i] = c1[i] + 0.2f;
        }
}
__device__ bool rgb_ordered(const float* c0, const float* c1, int q, int* c0i, int* c1i) {
    int c0q[4], c1q[4];
    const float f0[4] = {c0[0], c0[1], c0[2], 0.0f}, f1[4] = {c1[0], c1[1], c1[2], 0.0f};
    for (int i = 0; i < 3; ++i) {
        c0q[i] = max(wf2i(wround(c0[i])), 0);
        c1q[i] = min(wf2i(wround(c1[i])), 255);
    }
    c0q[3] = c1q[3] = 0;
    quant3_flt(q, c0q, f0, c0i);
    quant3_flt(q, c1q, f1, c1i);
    return (c0i[0] + c0i[1] + c0i[2]) <= (c1i[0] + c1i[1] + c1i[2]);
}
// kGallop = false keeps the plain step loop: the search costs registers
// (one wave per SIMD less for the 1-partition refinement, where the long
// loops are rare), and the 2..4-partition refinement, where they make the
// launch tail, is at its register budget either way.
template <bool kGallop>
__device__ void quantize_rgb(const float* color0, const float* color1, int* out0, int* out1, int q) {
    // the first steps one by one, as the reference takes them (most endpoints
    // are ordered at once or after a few steps)
    if (!kGallop) {
        float c0[4], c1[4];
        for (int i = 0; i < 4; ++i) { c0[i] = color0[i]; c1[i] = color1[i]; }
        int c0i[4], c1i[4];
        for (;;) {
            int c0q[4], c1q[4];
            for (int i = 0; i < 4; ++i) c0q[i] = max(wf2i(wround(c0[i])), 0);
            quant3_flt(q, c0q, c0, c0i);
            for (int i = 0; i < 4; ++i) c0[i] = c0[i] - 0.2f;
            for (int i = 0; i < 4; ++i) c1q[i] = min(wf2i(wround(c1[i])), 255);
            quant3_flt(q, c1q, c1, c1i);
            for (int i = 0; i < 4; ++i) c1[i] = c1[i] + 0.2f;
            if ((c0i[0] + c0i[1] + c0i[2]) <= (c1i[0] + c1i[1] + c1i[2])) break;
        }
        for (int i = 0; i < 4; ++i) { out0[i] = c0i[i]; out1[i] = c1i[i]; }
        return;
    }
    float l0[3] = {color0[0], color0[1], color0[2]}, l1[3] = {color1[0], color1[1], color1[2]};
    for (int s = 0;; ++s) {
        if (rgb_ordered(l0, l1, q, out0, out1)) return;
        if (s == 7) break;
        rgb_steps(l0, l1, 1);
    }
    // gallop: the state l (step lo) is unordered; probe lo + span, doubling span
    float m0[3], m1[3];
    int span = 1;
    for (;;) {
        for (int i = 0; i < 3; ++i) { m0[i] = l0[i]; m1[i] = l1[i]; }
        rgb_steps(m0, m1, span);
        if (rgb_ordered(m0, m1, q, out0, out1)) break;
        for (int i = 0; i < 3; ++i) { l0[i] = m0[i]; l1[i] = m1[i]; }
        if (span >= (1 << 14)) break;  // unreachable for finite endpoints (ordered within 1300 steps)
        span *= 2;
    }
    // bisect: lo unordered, lo + span ordered; then the quantised endpoints
    // of the first ordered step (lo + 1)
    int lo = 0, hi = span;
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        for (int i = 0; i < 3; ++i) { m0[i] = l0[i]; m1[i] = l1[i]; }
        rgb_steps(m0, m1, mid - lo);
        if (rgb_ordered(m0, m1, q, out0, out1)) {
            hi = mid;
        } else {
            lo = mid;
            for (int i = 0; i < 3; ++i) { l0[i] = m0[i]; l1[i] = m1[i]; }
        }
    }
    rgb_steps(l0, l1, 1);
    (void)rgb_ordered(l0, l1, q, out0, out1);
}
// try_quantize_alpha_delta, pass14:748-797
__device__ bool try_alpha_delta(const float* color0, const float* color1, int* out0, int* out1, int q) {
    const int a0a = wf2i(wround(color0[3])) << 1;
    int a0b = a0a & 0xFF;
    const int a0be = quant1_int(q, a0b);
    a0b = a0be | (a0a & 0x100);
    int a1d = (wf2i(wround(color1[3])) << 1) - a0b;
    if (a1d > 63 || a1d < -64) return false;
    a1d = (a1d & 0x7F) | ((a0b & 0x100) >> 1);
    const int a1de = quant1_int(q, a1d);
    int a1du = a1de;
    if (((a1d ^ a1du) & 0xC0) != 0) return false;
    a1du &= 0x7F;
    if (a1du & 0x40) a1du -= 0x80;
    a1du += a0b;
    if (a1du < 0 || a1du > 0x1FF) return false;
    out0[3] = a0be;
    out1[3] = a1de;
    return true;
}
// quantize_rgbs, pass14:858-884
__device__ void quantize_rgbs(const float* color, uint32_t* out, int q) {
    const float scale = 1.0f / 257.0f;
    const float r = wclamp(color[0] * scale, 0.0f, 255.0f);
    const float g = wclamp(color[1] * scale, 0.0f, 255.0f);
    const float b = wclamp(color[2] * scale, 0.0f, 255.0f);
    const int ri = quant1_flt(q, wf2i(wround(r)), r);
    const int gi = quant1_flt(q, wf2i(wround(g)), g);
    const int bi = quant1_flt(q, wf2i(wround(b)), b);
    const float oldsum = ((color[0] + color[1]) + color[2]) * scale;
    const float newsum = (float)(uint32_t)(ri + gi + bi);
    const float scalea = wclamp((color[3] * (oldsum + 1e-10f)) / (newsum + 1e-10f), 0.0f, 1.0f);
    const int sidx = min(max(wf2i(wround(scalea * 256.0f)), 0), 255);
    out[0] = (uint32_t)ri; out[1] = (uint32_t)gi; out[2] = (uint32_t)bi;
    out[3] = (uint32_t)quant1_int(q, sidx);
}
__device__ __forceinline__ float lum3(const float* c) { return ((c[0] + c[1]) + c[2]) / 3.0f; }
__device__ bool try_la_delta(const float* c0, const float* c1, uint32_t* out, int q) {
    const float l0 = lum3(c0), l1 = lum3(c1);
    const int l0a = wf2i(wround(l0)) << 1, a0a = wf2i(wround(c0[3])) << 1;
    int l0b = l0a & 0xFF, a0b = a0a & 0xFF;
    const int l0be = quant1_int(q, l0b), a0be = quant1_int(q, a0b);
    l0b = l0be | (l0a & 0x100);
    a0b = a0be | (a0a & 0x100);
    int l1d = (wf2i(wround(l1)) << 1) - l0b;
    int a1d = (wf2i(wround(c1[3])) << 1) - a0b;
    if (l1d > 63 || l1d < -64 || a1d > 63 || a1d < -64) return false;
    l1d = (l1d & 0x7F) | ((l0b & 0x100) >> 1);
    a1d = (a1d & 0x7F) | ((a0b & 0x100) >> 1);
    const int l1de = quant1_int(q, l1d), a1de = quant1_int(q, a1d);
    int l1du = l1de, a1du = a1de;
    if (((l1d ^ l1du) & 0xC0) != 0) return false;
    if (((a1d ^ a1du) & 0xC0) != 0) return false;
    l1du &= 0x7F; a1du &= 0x7F;
    if (l1du & 0x40) l1du -= 0x80;
    if (a1du & 0x40) a1du -= 0x80;
    l1du += l0b; a1du += a0b;
    if (l1du < 0 || l1du > 0x1FF || a1du < 0 || a1du > 0x1FF) return false;
    out[0] = (uint32_t)l0be; out[1] = (uint32_t)l1de; out[2] = (uint32_t)a0be; out[3] = (uint32_t)a1de;
    return true;
}

// pack_color_endpoints_helper, pass14:1018-1221
template <bool kGallop>
__device__ int pack_color_endpoints(const float* color0, const float* color1, const float* rgbs, int format, int q,
                                    uint32_t* packed) {
    float c0[4], c1[4];
    for (int i = 0; i < 4; ++i) {
        c0[i] = wclamp(color0[i], 0.0f, 65535.0f) / 257.0f;
        c1[i] = wclamp(color1[i], 0.0f, 65535.0f) / 257.0f;
    }
    float best_error = kErrDefault;
    int best_format = FMT_LUMINANCE;
    for (int i = 0; i < 8; ++i) packed[i] = 0;
    int o0[4] = {0, 0, 0, 0}, o1[4] = {0, 0, 0, 0}, t0[4] = {0, 0, 0, 0}, t1[4] = {0, 0, 0, 0}, u0[4], u1[4];
    auto take = [&](float e, int f) {
        if (e < best_error) {
            best_error = e; best_format = f;
            for (int i = 0; i < 4; ++i) { o0[i] = t0[i]; o1[i] = t1[i]; }
        }
    };
    switch (format) {
        case FMT_RGB:
        case FMT_RGBA: {
            const bool alpha = format == FMT_RGBA;
            if (q <= 18) {
                bool ok = try_rgb_delta_bc(c0, c1, o0, o1, q);
                if (ok && alpha) ok = try_alpha_delta(c1, c0, o0, o1, q);
                if (ok) {
                    rgba_delta_unpack(o0, o1, u0, u1);
                    best_format = alpha ? FMT_RGBA_DELTA : FMT_RGB_DELTA;
                    best_error = enc_error(c0, c1, u0, u1);
                }
                bool ok2 = try_rgb_delta_core(c0, c1, t0, t1, q, false);
                if (ok2 && alpha) ok2 = try_alpha_delta(c0, c1, t0, t1, q);
                if (ok2) {
                    rgba_delta_unpack(t0, t1, u0, u1);
                    take(enc_error(c0, c1, u0, u1), alpha ? FMT_RGBA_DELTA : FMT_RGB_DELTA);
                }
            }
            if (q < 20 && try_rgb_bc(c0, c1, t0, t1, q)) {
                if (alpha) {
                    t0[3] = quant1_flt(q, wf2i(wround(c1[3])), c1[3]);
                    t1[3] = quant1_flt(q, wf2i(wround(c0[3])), c0[3]);
                }
                rgba_unpack(t0, t1, u0, u1);
                take(enc_error(c0, c1, u0, u1), alpha ? FMT_RGBA : FMT_RGB);
            }
            quantize_rgb<kGallop>(c0, c1, t0, t1, q);
            if (alpha) {
                t0[3] = quant1_flt(q, wf2i(wround(c0[3])), c0[3]);
                t1[3] = quant1_flt(q, wf2i(wround(c1[3])), c1[3]);
            }
            rgba_unpack(t0, t1, u0, u1);
            take(enc_error(c0, c1, u0, u1), alpha ? FMT_RGBA : FMT_RGB);
            packed[0] = (uint32_t)o0[0]; packed[1] = (uint32_t)o1[0];
            packed[2] = (uint32_t)o0[1]; packed[3] = (uint32_t)o1[1];
            packed[4] = (uint32_t)o0[2]; packed[5] = (uint32_t)o1[2];
            if (alpha) { packed[6] = (uint32_t)o0[3]; packed[7] = (uint32_t)o1[3]; }
            break;
        }
        case FMT_RGB_SCALE:
            quantize_rgbs(rgbs, packed, q);
            best_format = FMT_RGB_SCALE;
            break;
        case FMT_RGB_SCALE_ALPHA:
            packed[4] = (uint32_t)quant1_flt(q, wf2i(wround(c0[3])), c0[3]);
            packed[5] = (uint32_t)quant1_flt(q, wf2i(wround(c1[3])), c1[3]);
            quantize_rgbs(rgbs, packed, q);
            best_format = FMT_RGB_SCALE_ALPHA;
            break;
        case FMT_LUMINANCE_ALPHA:
            if (q <= 18 && try_la_delta(c0, c1, packed, q)) { best_format = FMT_LUMINANCE_ALPHA_DELTA; break; }
            {
                const float l0 = lum3(c0), l1 = lum3(c1);
                packed[0] = (uint32_t)quant1_flt(q, wf2i(wround(l0)), l0);
                packed[1] = (uint32_t)quant1_flt(q, wf2i(wround(l1)), l1);
                packed[2] = (uint32_t)quant1_flt(q, wf2i(wround(c0[3])), c0[3]);
                packed[3] = (uint32_t)quant1_flt(q, wf2i(wround(c1[3])), c1[3]);
            }
            best_format = FMT_LUMINANCE_ALPHA;
            break;
        default: {  // FMT_LUMINANCE and anything else
            float l0 = lum3(c0), l1 = lum3(c1);
            if (l0 > l1) { const float avg = (l0 + l1) * 0.5f; l0 = avg; l1 = avg; }
            packed[0] = (uint32_t)quant1_flt(q, wf2i(wround(l0)), l0);
            packed[1] = (uint32_t)quant1_flt(q, wf2i(wround(l1)), l1);
            best_format = FMT_LUMINANCE;
            break;
        }
    }
    return best_format;
}

// LDR endpoint decode, pass15_unpack_color_endpoints.wgsl:307-391
__device__ void unpack_color_endpoints(int format, const uint32_t* v, int* o0, int* o1) {
    int in0[4], in1[4];
    switch (format) {
        case FMT_LUMINANCE_DELTA: {
            const int v0 = (int)v[0], v1 = (int)v[1];
            const int l0 = (v0 >> 2) | (v1 & 0xC0);
            const int l1 = min(l0 + (v1 & 0x3F), 255);
            o0[0] = o0[1] = o0[2] = l0; o0[3] = 255;
            o1[0] = o1[1] = o1[2] = l1; o1[3] = 255;
            return;
        }
        case FMT_LUMINANCE_ALPHA:
            o0[0] = o0[1] = o0[2] = (int)v[0]; o0[3] = (int)v[2];
            o1[0] = o1[1] = o1[2] = (int)v[1]; o1[3] = (int)v[3];
            return;
        case FMT_LUMINANCE_ALPHA_DELTA: {
            int l0 = (int)v[0], l1 = (int)v[1], a0 = (int)v[2], a1 = (int)v[3];
            l0 |= (l1 & 0x80) << 1;
            a0 |= (a1 & 0x80) << 1;
            l1 &= 0x7F; a1 &= 0x7F;
            if (l1 & 0x40) l1 -= 0x80;
            if (a1 & 0x40) a1 -= 0x80;
            l0 >>= 1; l1 >>= 1; a0 >>= 1; a1 >>= 1;
            l1 = min(max(l0 + l1, 0), 255);
            a1 = min(max(a0 + a1, 0), 255);
            o0[0] = o0[1] = o0[2] = l0; o0[3] = a0;
            o1[0] = o1[1] = o1[2] = l1; o1[3] = a1;
            return;
        }
        case FMT_RGB_SCALE: {
            const int s = (int)v[3];
            const int i0[4] = {(int)v[0], (int)v[1], (int)v[2], 255};
            for (int c = 0; c < 4; ++c) { o1[c] = i0[c]; o0[c] = (i0[c] * s) >> 8; }
            o0[3] = 255;
            return;
        }
        case FMT_RGB_SCALE_ALPHA: {
            const int s = (int)v[3];
            const int i0[4] = {(int)v[0], (int)v[1], (int)v[2], (int)v[4]};
            o1[0] = i0[0]; o1[1] = i0[1]; o1[2] = i0[2]; o1[3] = (int)v[5];
            for (int c = 0; c < 4; ++c) o0[c] = (i0[c] * s) >> 8;
            o0[3] = (int)v[4];
            return;
        }
        case FMT_RGB:
        case FMT_RGB_DELTA:
        case FMT_RGBA:
        case FMT_RGBA_DELTA: {
            const bool alpha = format == FMT_RGBA || format == FMT_RGBA_DELTA;
            in0[0] = (int)v[0]; in0[1] = (int)v[2]; in0[2] = (int)v[4]; in0[3] = alpha ? (int)v[6] : 0;
            in1[0] = (int)v[1]; in1[1] = (int)v[3]; in1[2] = (int)v[5]; in1[3] = alpha ? (int)v[7] : 0;
            if (format == FMT_RGB || format == FMT_RGBA) rgba_unpack(in0, in1, o0, o1);
            else rgba_delta_unpack(in0, in1, o0, o1);
            if (!alpha) { o0[3] = 255; o1[3] = 255; }
            return;
        }
        default:  // FMT_LUMINANCE and anything else
            o0[0] = o0[1] = o0[2] = (int)v[0]; o0[3] = 255;
            o1[0] = o1[1] = o1[2] = (int)v[1]; o1[3] = 255;
            return;
    }
}

// ---- passes 13-17, refinement loop of compression_pipeline.cpp:473-484 ---------
// One lane per (partitioned block, candidate slot).  The candidates are
// first bucketed by trial block mode and format class (k_top counts, k_bins
// lays out buckets padded to whole workgroups, k_scatter fills them), so a
// workgroup's lanes share the mode (loop bounds and decimation-table reads
// are uniform) and mostly take the same endpoint packing path.  Each
// lane keeps its block's texels (small footprints), partition map, quantised
// weights and unpacked endpoints in its own LDS rows; the per-partition sums
// of pass13 are accumulated partition by partition (each partition still
// sees its texels in index order, so the float results are those of the
// interleaved loop).

// LDS of one k_refine workgroup: staged texels, unpacked endpoints of the
// pc partitions, the texel partition map (pc > 1 only), candidate and best
// quantised weights.  A smaller 1-partition layout fits more workgroups per CU.
size_t refine_lds_bytes(const DevTables& T, int wstride, int pc) {
    size_t bytes = 2 * (size_t)pc * 64 * sizeof(uchar4) + (pc > 1 ? (size_t)T.tc * 64 : 0) + 2 * (size_t)wstride * 64;
    if (T.tc <= T.refine_stage_max_tc) bytes += (size_t)T.tc * 64 * sizeof(uint32_t);
    return bytes;
}

struct PackOut {
    int format;
    uint2 v;  // eight 0..255 values, little-endian bytes
};

template <bool kGallop>
__device__ __noinline__ PackOut pack_partition(float4 e0, float4 e1, float4 rgbs, int format, int q) {
    const float a[4] = {e0.x, e0.y, e0.z, e0.w}, b[4] = {e1.x, e1.y, e1.z, e1.w};
    const float r[4] = {rgbs.x, rgbs.y, rgbs.z, rgbs.w};
    uint32_t v[8];
    PackOut o;
    o.format = pack_color_endpoints<kGallop>(a, b, r, format, q, v);
    o.v.x = v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24);
    o.v.y = v[4] | (v[5] << 8) | (v[6] << 16) | (v[7] << 24);
    return o;
}

__device__ __forceinline__ uint32_t pk_byte(uint2 v, int i) { return ((i < 4 ? v.x : v.y) >> (8 * (i & 3))) & 0xFF; }

__device__ __forceinline__ int4 load_ue(const uchar4 u) { return make_int4(u.x * 257, u.y * 257, u.z * 257, u.w * 257); }

// ---- passes 13-17: refinement of one candidate ------------------------------
// A candidate is a partitioned block x slot refined with one trial block mode
// k.  Every lane of a wave refines a candidate of the same mode, so the
// mode's tables are wave-uniform; the candidate's state lives in registers
// and in the lane's LDS rows ([index][lane]).

struct RefineMode {
    int k;
    int q;              // weight quant level
    int wc;             // weight count
    bool undecimated;
    const uint8_t* t2wc;
    const uint8_t* t2wi;
    const float* t2ww;
    const uint8_t* w2tc;
    const uint16_t* w2to;
    const uint8_t* w2ti;
};

__device__ __forceinline__ RefineMode refine_mode(const DevTables& T, int k) {
    const BmInfo bmi = T.bm[k];
    const DmInfo di = T.dm[bmi.dm];
    RefineMode m;
    m.k = k;
    m.q = bmi.quant;
    m.wc = di.weight_count;
    m.undecimated = T.tc == m.wc;
    m.t2wc = T.t2w_count + (size_t)bmi.dm * T.tc;
    m.t2wi = T.t2w_idx + (size_t)bmi.dm * T.tc * 4;
    m.t2ww = T.t2w_w + (size_t)bmi.dm * T.tc * 4;
    m.w2tc = T.w2t_count + bmi.dm * 64;
    m.w2to = T.w2t_off + bmi.dm * 64;
    m.w2ti = T.w2t_idx + di.w2t_base;
    return m;
}

// The lane's LDS rows, each pointer already offset by the lane (element i of
// a row is at [i * 64]).
struct RefineRows {
    uint32_t* tex;  // [tc] RGBA8 texels (staged kernels)
    uchar4* ue0;    // [pc] unpacked endpoints
    uchar4* ue1;
    uint8_t* tp;    // [tc] texel partitions (pc > 1)
    uint8_t* qw;    // [wstride] quantised weights of the current round
    uint8_t* bqw;   // [wstride] quantised weights of the best round
    const float* u8f;  // [256] texel channel byte -> search value (u8_search_value), shared by the lanes
};

template <bool kStage, int PC>
__device__ __forceinline__ RefineRows refine_rows(uint8_t* smem, const DevTables& T, int wstride, int lane,
                                                  const float* u8f) {
    const int tc = T.tc;
    uint32_t* s_tex = reinterpret_cast<uint32_t*>(smem);
    uchar4* s_ue0 = reinterpret_cast<uchar4*>(smem + (kStage ? (size_t)tc * 64 * sizeof(uint32_t) : 0));
    uchar4* s_ue1 = s_ue0 + PC * 64;
    uint8_t* s_tp = reinterpret_cast<uint8_t*>(s_ue1 + PC * 64);
    uint8_t* s_qw = s_tp + (PC > 1 ? (size_t)tc * 64 : 0);
    uint8_t* s_bqw = s_qw + (size_t)wstride * 64;
    RefineRows r;
    r.tex = s_tex + lane;
    r.ue0 = s_ue0 + lane;
    r.ue1 = s_ue1 + lane;
    r.tp = s_tp + lane;
    r.qw = s_qw + lane;
    r.bqw = s_bqw + lane;
    r.u8f = u8f;
    return r;
}

struct RefineCand {
    int task;           // partitioned block * 8 + slot, -1 = none
    int ql, qlm;        // colour quant level and its mod variant (pass11)
    uint16_t in_formats;
    int ppc[4];         // texels per partition
    BlockLoc loc;         // texel origin in the image (unstaged 1-partition kernels read it in every pass)
    const uint32_t* tex;  // the block's texel row (unstaged partitioned kernels read it in every pass)
    float4 ep0[4], ep1[4];
    float sdirs[4][3];  // pass13's scaled direction: depends on the average only
    float best_err;
    int best_matched, best_quant;
    uint16_t best_formats;
    uint2 best_pk[4];
};

// pass13's sums over one partition's texels (weights, their complements,
// colour and scalar projections), taken in texel order.
struct Pass13Sums {
    float wmin1 = 1.0f, wmax1 = 0.0f, ls = 0.0f, ms = 0.0f, rs = 0.0f, smin = 1e10f, smax = 0.0f;
    float cx[4] = {0.0f, 0.0f, 0.0f, 0.0f}, cy[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float sv0 = 0.0f, sv1 = 0.0f;
    __device__ __forceinline__ void add(float w, const float* rgba, float sc, float lsw) {
        if (w < wmin1) wmin1 = w;
        if (w > wmax1) wmax1 = w;
        if (sc < smin) smin = sc;
        if (sc > smax) smax = sc;
        const float om = 1.0f - w;
        ls += om * om;
        ms += om * w;
        rs += w * w;
        for (int j = 0; j < 4; ++j) {
            const float cyk = rgba[j] * w;
            cy[j] += cyk;
            cx[j] += rgba[j] - cyk;
        }
        sv0 += om * sc * lsw;
        sv1 += w * sc * lsw;
    }
};

// Loads candidate `task` of mode M into the lane's rows and registers.
template <bool kStage, int PC>
__device__ __forceinline__ void refine_init(const ImageArgs& A, const DevTables& T, const int16_t* pb_part, int npb,
                                            const Ideal* ideal, const Ccr* ccr, const uint8_t* qw_all,
                                            const RefineMode& M, const RefineRows& R, int task, RefineCand& C) {
    constexpr int pc = PC;
    const int tc = T.tc;
    const int pb = task >> 3;
    const Pb b = make_pb(A, T, pb_part, pc, pb);
    for (int i = 0; i < tc; ++i) {
        if (kStage) R.tex[i * 64] = A.blk_texels[(size_t)(b.g - A.blk0) * tc + i];
        if (pc > 1) R.tp[i * 64] = (uint8_t)texel_part(b, i);
    }
    C.task = task;
    C.loc = b.loc;
    C.tex = A.blk_texels + (size_t)(b.g - A.blk0) * tc;
#pragma unroll
    for (int i = 0; i < 4; ++i) C.ppc[i] = b.ppc[i];
    const Ideal& I = ideal[pb];
    const Ccr cr = ccr[(size_t)M.k * npb + pb];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        C.ep0[p] = make_float4(I.ep0[p][0], I.ep0[p][1], I.ep0[p][2], I.ep0[p][3]);
        C.ep1[p] = make_float4(I.ep1[p][0], I.ep1[p][1], I.ep1[p][2], I.ep1[p][3]);
        const float size = (float)C.ppc[p];
        float q3[3];
        for (int j = 0; j < 3; ++j) q3[j] = ((I.avg[p][j] * size) * kCW[j]) / wmax(kCW[j] * size, 1e-17f);
        normalize3(q3, C.sdirs[p]);
    }
    C.ql = cr.ql;
    C.qlm = cr.ql_mod;
    C.in_formats = cr.formats;
    {
        // this PB's row of the mode's region, read as dwords
        const uint32_t* qsrc = reinterpret_cast<const uint32_t*>(qw_all + (size_t)T.qw_off[M.k] * npb +
                                                                 (size_t)pb * ((M.wc + 3) & ~3));
        for (int i = 0; i < M.wc; i += 4) {
            const uint32_t v = qsrc[i >> 2];
            for (int j = 0; j < 4 && i + j < M.wc; ++j) R.qw[(i + j) * 64] = (uint8_t)(v >> (8 * j));
        }
    }
    C.best_err = kErrDefault;
    C.best_matched = 0;
    C.best_quant = 0;
    C.best_formats = 0;
#pragma unroll
    for (int p = 0; p < 4; ++p) C.best_pk[p] = make_uint2(0, 0);
}

// One refinement round (pass13, pass14, pass15, then pass17 on the first
// round and pass16 + pass17); returns whether pass16 moved a weight.  Once
// pass16 leaves every weight where it was, the next round would recompute
// the same endpoints from the same weights (pass13 keeps a previous endpoint
// only where it kept it this time), so it would repeat this round's error and
// could not replace the best candidate: the caller stops there.
template <bool kStage, int PC>
__device__ __forceinline__ bool refine_round(const DevTables& T, const RefineMode& M, const RefineRows& R, bool first,
                                             RefineCand& C) {
    constexpr int pc = PC;
    const int tc = T.tc;
    // a texel's channels as search values, through the workgroup's table
    // (four LDS reads instead of the conversion's arithmetic)
    auto px_of = [&](int i) -> float4 {
        // one partition unstaged: neighbouring lanes' blocks share the
        // image's cache lines (measured faster than the block rows)
        const uint32_t v = kStage ? R.tex[i * 64] : PC == 1 ? texel_rgba8(C.loc, T, i) : C.tex[i];
        return make_float4(R.u8f[v & 0xFF], R.u8f[(v >> 8) & 0xFF], R.u8f[(v >> 16) & 0xFF], R.u8f[v >> 24]);
    };
    auto part_of = [&](int i) -> int { return pc > 1 ? (int)R.tp[i * 64] : 0; };
    uint8_t* qw = R.qw;
    const int wc = M.wc;
    const bool undecimated = M.undecimated;
    const uint8_t* t2wc = M.t2wc;
    const uint8_t* t2wi = M.t2wi;
    const float* t2ww = M.t2ww;
    const float lsw = (kCW[0] + kCW[1]) + kCW[2];
    const float cw[4] = {kCW[0], kCW[1], kCW[2], kCW[3]};
    auto weight_of = [&](int i) -> float {
        if (undecimated) return (float)qw[i * 64] / 64.0f;
        float w = 0.0f;
        const int n = t2wc[i];
        for (int j = 0; j < n; ++j) w += ((float)qw[t2wi[i * 4 + j] * 64] / 64.0f) * t2ww[i * 4 + j];
        return w;
    };
    // pass13_recompute_ideal_endpoints.wgsl:182-393
    // Per-partition sums over the partition's texels in texel order.  With
    // two partitions one walk over the texels fills both sets (each set
    // still receives its texels in index order), instead of one divergent
    // walk per partition.
    Pass13Sums sums2[2];
    if (pc == 2) {
        for (int i = 0; i < tc; ++i) {
            const int p = R.tp[i * 64];
            const float w = weight_of(i);
            const float4 px4 = px_of(i);
            const float rgba[4] = {px4.x, px4.y, px4.z, px4.w};
            const float* sd = p ? C.sdirs[1] : C.sdirs[0];
            const float sc = dot3(rgba, sd);
            if (p) sums2[1].add(w, rgba, sc, lsw);
            else sums2[0].add(w, rgba, sc, lsw);
        }
    }
    float4 rgbs[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        rgbs[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (p >= pc) continue;
        const float size = (float)C.ppc[p];
        const float* sdir = C.sdirs[p];
        Pass13Sums acc;
        if (pc == 2) {
            acc = sums2[p & 1];
        } else {
            for (int i = 0; i < tc; ++i) {
                if (pc > 1 && R.tp[i * 64] != p) continue;
                const float w = weight_of(i);
                const float4 px4 = px_of(i);
                const float rgba[4] = {px4.x, px4.y, px4.z, px4.w};
                acc.add(w, rgba, dot3(rgba, sdir), lsw);
            }
        }
        const float wmin1 = acc.wmin1, wmax1 = acc.wmax1, ls = acc.ls, ms = acc.ms, rs = acc.rs;
        const float smin = acc.smin, smax = acc.smax, sv0 = acc.sv0, sv1 = acc.sv1;
        const float* cx = acc.cx;
        const float* cy = acc.cy;
        float colx[4], coly[4], lv[4], mv[4], rv[4];
        for (int j = 0; j < 4; ++j) {
            colx[j] = cx[j] * cw[j]; coly[j] = cy[j] * cw[j];
            lv[j] = ls * cw[j]; mv[j] = ms * cw[j]; rv[j] = rs * cw[j];
        }
        const float lx = ls * lsw, ly = ms * lsw, lz = rs * lsw;
        const float scalediv = wclamp(smin / wmax(smax, 1e-10f), 0.0f, 1.0f);
        rgbs[p] = make_float4(sdir[0] * smax, sdir[1] * smax, sdir[2] * smax, scalediv);
        float e0v[4] = {C.ep0[p].x, C.ep0[p].y, C.ep0[p].z, C.ep0[p].w};
        float e1v[4] = {C.ep1[p].x, C.ep1[p].y, C.ep1[p].z, C.ep1[p].w};
        if (wmin1 >= wmax1 * 0.999f) {
            for (int j = 0; j < 4; ++j) {
                const float ac = (colx[j] + coly[j]) / wmax(cw[j] * size, 1e-17f);
                if (ac == ac) { e0v[j] = ac; e1v[j] = ac; }
            }
            rgbs[p].w = 1.0f;
        } else {
            for (int j = 0; j < 4; ++j) {
                const float det = (lv[j] * rv[j]) - (mv[j] * mv[j]);
                const float rdet = 1.0f / det;
                const float mss = ((lv[j] * lv[j]) + 2.0f * (mv[j] * mv[j])) + (rv[j] * rv[j]);
                const float e0 = (rv[j] * colx[j] - mv[j] * coly[j]) * rdet;
                const float e1 = (lv[j] * coly[j] - mv[j] * colx[j]) * rdet;
                if (fabsf(det) > (mss * 1e-4f) && e0 == e0 && e1 == e1) { e0v[j] = e0; e1v[j] = e1; }
            }
            const float lsdet = (lx * lz) - (ly * ly);
            const float lsrdet = 1.0f / lsdet;
            const float lsmss = ((lx * lx) + 2.0f * (ly * ly)) + (lz * lz);
            const float se0 = (lz * sv0 - ly * sv1) * lsrdet;
            const float se1 = (lx * sv1 - ly * sv0) * lsrdet;
            if ((fabsf(lsdet) > (lsmss * 1e-4f)) && se0 == se0 && se1 == se1 && se0 < se1)
                rgbs[p] = make_float4(sdir[0] * se1, sdir[1] * se1, sdir[2] * se1, se0 / se1);
        }
        C.ep0[p] = make_float4(e0v[0], e0v[1], e0v[2], e0v[3]);
        C.ep1[p] = make_float4(e1v[0], e1v[1], e1v[2], e1v[3]);
    }
    // pass14_pack_color_endpoints.wgsl:1226-1304
    int fn[4] = {0, 0, 0, 0};
    uint2 pk[4] = {make_uint2(0, 0), make_uint2(0, 0), make_uint2(0, 0), make_uint2(0, 0)};
    bool all_same = C.ql != C.qlm;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        if (p >= pc) continue;
        const PackOut o = pack_partition<(PC > 1)>(C.ep0[p], C.ep1[p], rgbs[p], fmt_of(C.in_formats, p), C.ql);
        fn[p] = o.format;
        pk[p] = o.v;
        if (p > 0) all_same = all_same && (o.format == fn[0]);
    }
    bool matched = false;
    if (pc >= 2 && all_same) {
        int fm[4] = {0, 0, 0, 0};
        uint2 pm[4] = {make_uint2(0, 0), make_uint2(0, 0), make_uint2(0, 0), make_uint2(0, 0)};
        bool same_mod = true;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            if (p >= pc || !same_mod) continue;
            const PackOut o = pack_partition<(PC > 1)>(C.ep0[p], C.ep1[p], rgbs[p], fmt_of(C.in_formats, p), C.qlm);
            fm[p] = o.format;
            pm[p] = o.v;
            if (p > 0 && o.format != fm[0]) same_mod = false;
        }
        if (same_mod) {
            matched = true;
#pragma unroll
            for (int p = 0; p < 4; ++p) { fn[p] = fm[p]; pk[p] = pm[p]; }
        }
    }
    const int final_quant = matched ? C.qlm : C.ql;
    int he0[4] = {0, 0, 0, 0}, he1[4] = {0, 0, 0, 0};  // 1 partition: decoded endpoints (x 257)
    // pass15_unpack_color_endpoints.wgsl:307-391
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        if (p >= pc) continue;
        uint32_t v[8];
        for (int i = 0; i < 8; ++i) v[i] = pk_byte(pk[p], i);
        int o0[4], o1[4];
        unpack_color_endpoints(fn[p], v, o0, o1);
        if (PC == 1) {
            for (int c = 0; c < 4; ++c) {
                he0[c] = (int)(uint8_t)o0[c] * 257;
                he1[c] = (int)(uint8_t)o1[c] * 257;
            }
        } else {
            R.ue0[p * 64] = make_uchar4(o0[0], o0[1], o0[2], o0[3]);
            R.ue1[p * 64] = make_uchar4(o1[0], o1[1], o1[2], o1[3]);
        }
    }
    // the endpoints of texel t's partition as the texel loops use them: the
    // interpolation base (decoded endpoint 0, x 257) and step per channel;
    // with one or two partitions they are formed once per round and stay
    // in registers, with more they are decoded per texel from LDS
    float pbs[2][4], pof[2][4];
    if (PC <= 2) {
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            if (p >= pc) continue;
            int4 u0, u1;
            if (PC == 1) {
                u0 = make_int4(he0[0], he0[1], he0[2], he0[3]);
                u1 = make_int4(he1[0], he1[1], he1[2], he1[3]);
            } else {
                u0 = load_ue(R.ue0[p * 64]);
                u1 = load_ue(R.ue1[p * 64]);
            }
            const int a0[4] = {u0.x, u0.y, u0.z, u0.w}, a1[4] = {u1.x, u1.y, u1.z, u1.w};
            for (int j = 0; j < 4; ++j) {
                pbs[p][j] = (float)a0[j];
                pof[p][j] = (float)(a1[j] - a0[j]) * (1.0f / 64.0f);
            }
        }
    }
    // (e0, e1: the decoded endpoints, kept as integers where they are
    // decoded per texel)
    auto ep_at = [&](int t, float* bs, float* of, int* e0, int* e1) {
        if (PC <= 2) {
            const bool p1 = PC == 2 && part_of(t) != 0;
            for (int j = 0; j < 4; ++j) {
                bs[j] = p1 ? pbs[1][j] : pbs[0][j];
                of[j] = p1 ? pof[1][j] : pof[0][j];
            }
        } else {
            const int p = part_of(t);
            const int4 u0 = load_ue(R.ue0[p * 64]), u1 = load_ue(R.ue1[p * 64]);
            e0[0] = u0.x; e0[1] = u0.y; e0[2] = u0.z; e0[3] = u0.w;
            e1[0] = u1.x; e1[1] = u1.y; e1[2] = u1.z; e1[3] = u1.w;
            for (int j = 0; j < 4; ++j) {
                bs[j] = (float)e0[j];
                of[j] = (float)(e1[j] - e0[j]) * (1.0f / 64.0f);
            }
        }
    };
    // pass17's reconstruction (e0 (64 - w) + e1 w + 32) >> 6 of channel j.
    // From registers it runs in float on e0 = bs, e1 = bs + 64 of: every term
    // is an integer below 2^24 (e <= 65535, w <= 64) and / 64 is a power of
    // two, so each step is exact and the floor is the shift.
    auto recon = [&](const float* bs, const float* of, const int* e0, const int* e1, int j, int w0, int w1) {
        if (PC <= 2) {
            const float e1f = bs[j] + of[j] * 64.0f;
            return floorf(((bs[j] * (float)w0 + e1f * (float)w1) + 32.0f) * (1.0f / 64.0f));
        }
        return (float)((e0[j] * w0 + e1[j] * w1 + 32) >> 6);
    };
    // pass17 (before pass16 on the first round), then pass16 + pass17
    bool moved = false;
    for (int phase = (first ? 0 : 1); phase < 2; ++phase) {
        float total = 0.0f;  // pass17's error sum
        bool fused = false;  // pass17 already summed inside pass16's texel loop
        if (phase == 1) {
            // pass16_realign_weights.wgsl:211-367
            const int q = M.q;
            if (undecimated) {
                // every texel owns its weight, so texel t's final error
                // (pass17) depends only on its own realigned weight: both
                // passes share one walk over the texels, in texel order
                fused = true;
                for (int t = 0; t < tc; ++t) {
                    const int uqw = qw[t * 64];
                    const uint32_t pn = c_qt.prev_next[q][uqw];
                    const int down = (int)(pn & 0xFF), up = (int)((pn >> 8) & 0xFF);
                    const float wb = (float)uqw, dd = (float)(down - uqw), udf = (float)(up - uqw);
                    float bs[4], of[4];
                    int e0[4], e1[4];
                    ep_at(t, bs, of, e0, e1);
                    const float4 px4 = px_of(t);
                    const float px[4] = {px4.x, px4.y, px4.z, px4.w};
                    float s0[4], s1[4], s2[4];
                    for (int j = 0; j < 4; ++j) {
                        const float base = bs[j], offs = of[j];
                        const float diff = (base + offs * wb) - px[j];
                        const float dfd = diff + offs * dd, dfu = diff + offs * udf;
                        s0[j] = diff * diff; s1[j] = dfd * dfd; s2[j] = dfu * dfu;
                    }
                    const float eb = dot4(s0, cw), ed = dot4(s1, cw), eu = dot4(s2, cw);
                    int nq = uqw;
                    if ((eu < eb) && (eu < ed) && (uqw < 64)) nq = up;
                    else if ((ed < eb) && (uqw > 0)) nq = down;
                    moved |= nq != uqw;
                    qw[t * 64] = (uint8_t)nq;
                    // pass17_compute_final_error.wgsl:149-231 for this texel
                    float dfe[4];
                    for (int j = 0; j < 4; ++j) {
                        const float col = recon(bs, of, e0, e1, j, 64 - nq, nq);
                        const float df = wmin(fabsf(px[j] - col), 1e15f);
                        dfe[j] = df * df;
                    }
                    total += wmin(dot4(dfe, cw), kErrDefault);
                }
            } else {
                const uint8_t* w2tc = M.w2tc;
                const uint16_t* w2to = M.w2to;
                const uint8_t* w2ti = M.w2ti;
                for (int we = 0; we < wc; ++we) {
                    const int uqw = qw[we * 64];
                    const uint32_t pn = c_qt.prev_next[q][uqw];
                    const float down = (float)(pn & 0xFF), up = (float)((pn >> 8) & 0xFF), wbase = (float)uqw;
                    const float ddn = down - wbase, dup = up - wbase;
                    float ebv[4] = {0, 0, 0, 0}, edv[4] = {0, 0, 0, 0}, euv[4] = {0, 0, 0, 0};
                    const int n = w2tc[we], woff = w2to[we];
                    for (int te = 0; te < n; ++te) {
                        const int t = w2ti[woff + te];
                        float wbt = 0.0f, twb = 0.0f;
                        bool found = false;
                        for (int j = 0; j < t2wc[t]; ++j) {
                            const int wi = t2wi[t * 4 + j];
                            wbt += (float)qw[wi * 64] * t2ww[t * 4 + j];
                            if (!found && wi == we) { twb = t2ww[t * 4 + j]; found = true; }
                        }
                        const float wdd = ddn * twb, wud = dup * twb;
                        float bs[4], of[4];
                        int e0[4], e1[4];
                        ep_at(t, bs, of, e0, e1);
                        const float4 px4 = px_of(t);
                        const float px[4] = {px4.x, px4.y, px4.z, px4.w};
                        for (int j = 0; j < 4; ++j) {
                            const float base = bs[j], offs = of[j];
                            const float diff = (base + offs * wbt) - px[j];
                            const float dd = diff + offs * wdd, du = diff + offs * wud;
                            ebv[j] += diff * diff;
                            edv[j] += dd * dd;
                            euv[j] += du * du;
                        }
                    }
                    const float eb = dot4(ebv, cw), ed = dot4(edv, cw), eu = dot4(euv, cw);
                    int nq = uqw;
                    if ((eu < eb) && (eu < ed) && (uqw < 64)) nq = (int)up;
                    else if ((ed < eb) && (uqw > 0)) nq = (int)down;
                    moved |= nq != uqw;
                    qw[we * 64] = (uint8_t)nq;
                }
            }
        }
        // pass17_compute_final_error.wgsl:149-231
        for (int i = 0; i < tc && !fused; ++i) {
            const float w = weight_of(i);
            float bs[4], of[4];
            int e0[4], e1[4];
            ep_at(i, bs, of, e0, e1);
            const float4 px4 = px_of(i);
            const float px[4] = {px4.x, px4.y, px4.z, px4.w};
            const int w1 = wf2i(wround(w * 64.0f)), w0 = 64 - w1;
            float dd[4];
            for (int j = 0; j < 4; ++j) {
                const float col = recon(bs, of, e0, e1, j, w0, w1);
                const float df = wmin(fabsf(px[j] - col), 1e15f);
                dd[j] = df * df;
            }
            total += wmin(dot4(dd, cw), kErrDefault);
        }
        if (total < C.best_err) {
            C.best_err = total;
            C.best_matched = matched ? 1 : 0;
            C.best_quant = final_quant;
            C.best_formats = pack_fmt4(fn);
#pragma unroll
            for (int p = 0; p < 4; ++p) C.best_pk[p] = pk[p];
            for (int i = 0; i < wc; ++i) R.bqw[i * 64] = qw[i * 64];
        }
    }
    return moved;
}

// Writes the candidate's best round to its output slot.
__device__ __forceinline__ void refine_store(const RefineMode& M, const RefineRows& R, const RefineCand& C,
                                             CandOut* cand_out, float* cand_err) {
    CandOut best;
    best.error = C.best_err;
    best.bm = (uint16_t)M.k;
    best.matched = (uint8_t)C.best_matched;
    best.quant = (uint8_t)C.best_quant;
    best.formats = C.best_formats;
    best.pad = 0;
#pragma unroll
    for (int p = 0; p < 4; ++p)
        for (int i = 0; i < 8; ++i) best.packed[p * 8 + i] = (uint8_t)pk_byte(C.best_pk[p], i);
    for (int i = 0; i < 64; ++i) best.qw[i] = i < M.wc ? R.bqw[i * 64] : 0;
    cand_out[C.task] = best;
    cand_err[C.task] = best.error;
}

// One candidate per lane: a workgroup takes one 64-entry tile of the
// mode-bucketed work list (buckets padded to 64 lanes).  Specialised per
// partition count, so per-partition state only occupies registers for the
// partitions that exist.
template <bool kStage, int PC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PC == 1 ? 6 : 4, 8)))
k_refine(ImageArgs A, DevTables T, const int16_t* pb_part, int npb, const Ideal* ideal, const int16_t* top,
         const Ccr* ccr, const uint8_t* qw_all, const int* work, const int* work_total, CandOut* cand_out,
         float* cand_err, int wstride, int iterations) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = threadIdx.x;
    const int base = blockIdx.x * 64;
    if (base >= *work_total) return;
    __shared__ float s_u8f[256];
    for (int e = lane; e < 256; e += 64) s_u8f[e] = u8_search_value((uint32_t)e);
    __syncthreads();
    const int task = work[base + lane];
    const int k = __builtin_amdgcn_readfirstlane((int)top[__builtin_amdgcn_readfirstlane(work[base])]) >>
                  kRefineClassBits;
    if (task < 0) return;
    const RefineRows R = refine_rows<kStage, PC>(smem, T, wstride, lane, s_u8f);
    const RefineMode M = refine_mode(T, k);
    RefineCand C;
    refine_init<kStage, PC>(A, T, pb_part, npb, ideal, ccr, qw_all, M, R, task, C);
    for (int a = 0; a < iterations; ++a)  // 6 in the reference (pass13-17 rounds)
        if (!refine_round<kStage, PC>(T, M, R, a == 0, C)) break;
    refine_store(M, R, C, cand_out, cand_err);
}

// Persistent form: a grid of resident waves that take the work list's
// 64-entry tiles from a queue (tile_next, zeroed by k_bins) until it is
// empty.  A lane that finishes its candidate (pass16 moved no weight, or the
// last round) takes the next entry of the wave's current tile, so the lanes
// of a wave do not wait for its slowest candidate; a tile of another mode is
// held until every lane has finished (the wave's tables are one mode's).
// Free lanes are refilled once at least refill_min of them are free (each
// refill runs the candidate set-up with the other lanes masked off).  Every
// candidate's arithmetic is that of k_refine, and it writes the same output
// slot, so the results are identical.
template <bool kStage, int PC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PC == 1 ? 6 : 4, 8)))
k_refine_q(ImageArgs A, DevTables T, const int16_t* pb_part, int npb, const Ideal* ideal, const int16_t* top,
           const Ccr* ccr, const uint8_t* qw_all, const int* work, const int* work_total, int* tile_next,
           CandOut* cand_out, float* cand_err, int wstride, int iterations, int refill_min) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = threadIdx.x;
    const int ntiles = __builtin_amdgcn_readfirstlane(*work_total) >> 6;
    __shared__ float s_u8f[256];
    for (int e = lane; e < 256; e += 64) s_u8f[e] = u8_search_value((uint32_t)e);
    __syncthreads();
    const RefineRows R = refine_rows<kStage, PC>(smem, T, wstride, lane, s_u8f);
    RefineMode M{};
    int k = -1;                     // the wave's mode, -1 before its first tile
    int pool = 0, pool_end = 0;     // entries of the current tile not handed out yet
    int pending = -1;               // a tile of another mode, started after the drain
    bool more = true;               // the queue may hold tiles
    RefineCand C;
    C.task = -1;
    int a = 0;                      // the lane's round
    for (;;) {
        uint64_t fm = __ballot(C.task < 0);
        int nf = __popcll(fm);
        if (nf >= refill_min || nf == 64) {
            while (nf > 0) {
                if (pool >= pool_end) {
                    if (pending >= 0) {
                        if (nf < 64) break;  // lanes of the current mode still run
                        k = __builtin_amdgcn_readfirstlane((int)top[work[pending * 64]]) >> kRefineClassBits;
                        M = refine_mode(T, k);
                        pool = pending * 64;
                        pool_end = pool + 64;
                        pending = -1;
                    } else if (more) {
                        int t = 0;
                        if (lane == 0) t = atomicAdd(tile_next, 1);
                        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
                        if (t >= ntiles) {
                            more = false;
                            continue;
                        }
                        const int m = __builtin_amdgcn_readfirstlane((int)top[work[t * 64]]) >> kRefineClassBits;
                        if (k >= 0 && m != k) {
                            pending = t;
                            continue;
                        }
                        if (k < 0) {
                            k = m;
                            M = refine_mode(T, k);
                        }
                        pool = t * 64;
                        pool_end = pool + 64;
                    } else {
                        break;
                    }
                }
                // free lanes take the tile's next entries in lane order;
                // padding (-1) ends a bucket's last tile
                const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(fm >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)fm, 0u));
                const int avail = pool_end - pool;
                const bool mine = C.task < 0 && rank < avail;
                const int tk = mine ? work[pool + rank] : -1;
                pool += nf < avail ? nf : avail;
                if (__ballot(mine && tk < 0)) pool = pool_end;
                if (tk >= 0) {
                    refine_init<kStage, PC>(A, T, pb_part, npb, ideal, ccr, qw_all, M, R, tk, C);
                    a = 0;
                }
                fm = __ballot(C.task < 0);
                nf = __popcll(fm);
            }
        }
        if (nf == 64) break;  // nothing left: a full refill found no work
        if (C.task >= 0) {
            const bool moved = refine_round<kStage, PC>(T, M, R, a == 0, C);
            ++a;
            if (!moved || a >= iterations) {
                refine_store(M, R, C, cand_out, cand_err);
                C.task = -1;
            }
        }
    }
}

// ---- pass18 + host argmin (compression_pipeline.cpp:499-519) --------------------
__global__ void __launch_bounds__(64) k_pick(ImageArgs A, DevTables T, int pc, const int16_t* pb_part,
                                             const CandOut* cand, const float* cand_err, SymBest* best,
                                             LabelBufs lab) {
    // one lane per block picks; the wave then copies its blocks' chosen
    // 96-byte packed / weight runs together (consecutive lanes on consecutive
    // dwords instead of one 96-byte stride per lane)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = k < A.nb;
    const int npos = pc == 1 ? 1 : 4;
    float cur_err = 0.0f;
    int cur_pb = -1, cur_slot = -1;
    for (int c = 0; c < (live ? npos : 0); ++c) {
        const int pb = (pc > 1 && A.pb_inv) ? A.pb_inv[k * 4 + c] : k * npos + c;
        if (pb < 0) {  // position not searched (learned policy)
            if (lab.cand_seed) {
                const size_t li = ((size_t)(pc - 2) * 4 + c) * A.num_blocks + A.blk0 + k;
                lab.cand_seed[li] = -1;
                lab.cand_err[li] = 1e30f;
            }
            continue;
        }
        float be = kErrDefault;
        int bs = -1;
        for (int s = 0; s < 8; ++s) {
            const float e = cand_err[pb * 8 + s];
            if (e < be) { be = e; bs = s; }
        }
        if (lab.cand_seed && pc > 1) {
            const size_t li = ((size_t)(pc - 2) * 4 + c) * A.num_blocks + A.blk0 + k;
            lab.cand_seed[li] = (int16_t)T.part_seed[(size_t)(pc - 2) * kPartitionings + pb_part[pb_slot_of(A, pc, pb)]];
            lab.cand_err[li] = bs >= 0 ? be : 1e30f;
        }
        if (cur_pb < 0 || be < cur_err) { cur_err = be; cur_pb = pb; cur_slot = bs; }
    }
    int src = -1;  // candidate whose packed / weight run this block takes
    if (live) {
        SymBest& sb = best[A.blk0 + k];
        // optional record of each partition count's best partitioning
        // (labels for the candidate predictor)
        if (lab.best_seed && pc > 1)
            lab.best_seed[(size_t)(pc - 2) * A.num_blocks + A.blk0 + k] =
                cur_slot >= 0
                    ? (int16_t)T.part_seed[(size_t)(pc - 2) * kPartitionings + pb_part[pb_slot_of(A, pc, cur_pb)]]
                    : (int16_t)-1;
        if (pc == 1) {
            sb.error = 1e30f;
            sb.raw_mode = 0xFFFF;
            sb.pc = 0;
            sb.pad = 0;  // single plane
        }
        if (cur_slot >= 0 && cur_err < sb.error) {
            src = cur_pb * 8 + cur_slot;
            const CandOut& co = cand[src];
            sb.error = cur_err;
            sb.raw_mode = T.bm[co.bm].raw_mode;
            sb.pc = (uint8_t)pc;
            int kb, cb, slot;
            pb_decode(A, pc, cur_pb, &kb, &cb, &slot);
            sb.seed = pc == 1 ? 0 : T.part_seed[(size_t)(pc - 2) * kPartitionings + pb_part[slot]];
            sb.matched = co.matched;
            sb.quant = co.quant;
            sb.formats = co.formats;
        }
    }
    // 64 blocks x 96 bytes as 24 dwords each (both records keep the arrays
    // 4-byte aligned); lane l of round i copies dword (64 i + l) % 24 of block
    // (64 i + l) / 24 of the wave
    const int wbase = blockIdx.x * blockDim.x;
    for (int i = 0; i < 24; ++i) {
        const int e = i * 64 + (int)threadIdx.x, b = e / 24, d = e - b * 24;
        const int so = __shfl(src, b);
        if (so >= 0)
            reinterpret_cast<uint32_t*>(best[A.blk0 + wbase + b].packed)[d] =
                reinterpret_cast<const uint32_t*>(cand[so].packed)[d];
    }
}

// ---- symbolic_to_physical (physical_compression.cpp:843-1003) ---------------------
__device__ __forceinline__ void write_bits(uint32_t value, int bitcount, int bitoffset, uint32_t* w) {
    // 128-bit block as four little-endian words; a field never spans more than two words
    const uint32_t mask = bitcount >= 32 ? 0xFFFFFFFFu : ((1u << bitcount) - 1u);
    value &= mask;
    const int wi = bitoffset >> 5, sh = bitoffset & 31;
    w[wi] = (w[wi] & ~(mask << sh)) | (value << sh);
    if (sh + bitcount > 32) {
        const int s2 = 32 - sh;
        w[wi + 1] = (w[wi + 1] & ~(mask >> s2)) | (value >> s2);
    }
}

__device__ void encode_ise(int q, int count, const uint8_t* in, uint32_t* w, int off) {
    int bits, trits = 0, quints = 0;
    {
        const int n = kLevelsD[q];
        trits = n % 3 == 0;
        quints = n % 5 == 0;
        const int base = trits ? n / 3 : (quints ? n / 5 : n);
        bits = 31 - __clz(base);
    }
    const uint32_t mask = (1u << bits) - 1u;
    if (trits) {
        const int tb[5] = {2, 2, 1, 2, 1}, ts[5] = {0, 2, 4, 5, 7};
        for (int i = 0; i < count; i += 5) {
            int t[5] = {0, 0, 0, 0, 0};
            for (int k = 0; k < 5; ++k) t[k] = (i + k < count) ? (in[i + k] >> bits) : 0;
            const int T = c_qt.int_of_trits[(((t[4] * 3 + t[3]) * 3 + t[2]) * 3 + t[1]) * 3 + t[0]];
            for (int k = 0; k < 5 && i + k < count; ++k) {
                const uint32_t pack = (in[i + k] & mask) | (((uint32_t)(T >> ts[k]) & ((1u << tb[k]) - 1u)) << bits);
                write_bits(pack, bits + tb[k], off, w);
                off += bits + tb[k];
            }
        }
    } else if (quints) {
        const int qb[3] = {3, 2, 2}, qs[3] = {0, 3, 5};
        for (int i = 0; i < count; i += 3) {
            int t[3] = {0, 0, 0};
            for (int k = 0; k < 3; ++k) t[k] = (i + k < count) ? (in[i + k] >> bits) : 0;
            const int Q = c_qt.int_of_quints[(t[2] * 5 + t[1]) * 5 + t[0]];
            for (int k = 0; k < 3 && i + k < count; ++k) {
                const uint32_t pack = (in[i + k] & mask) | (((uint32_t)(Q >> qs[k]) & ((1u << qb[k]) - 1u)) << bits);
                write_bits(pack, bits + qb[k], off, w);
                off += bits + qb[k];
            }
        }
    } else {
        for (int i = 0; i < count; ++i) {
            write_bits(in[i], bits, off, w);
            off += bits;
        }
    }
}

__device__ __forceinline__ uint32_t bitrev32(uint32_t v) { return __brev(v); }

__global__ void __launch_bounds__(64) k_pack(DevTables T, ImageArgs A, const SymBest* best, int void_extent) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.num_blocks) return;
    const SymBest& sb = best[k];
    const TexDesc t = tex_of(A, k);
    uint4* o = reinterpret_cast<uint4*>(t.out + (size_t)(k - t.first_block) * 16);
    if (void_extent) {
        // a block whose texels (edge-clamped) are all equal becomes an LDR
        // void-extent block (ASTC specification): block mode 0x1FC, LDR,
        // the two reserved bits set, all-ones (absent) extent coordinates,
        // then R, G, B, A as UNORM16 (u8 * 257, which decodes back exactly)
        const BlockLoc l = block_loc(A, T, k);
        const uint32_t c0 = texel_rgba8(l, T, 0);
        bool same = true;
        for (int i = 1; i < T.tc && same; ++i) same = texel_rgba8(l, T, i) == c0;
        if (same) {
            const uint32_t r = c0 & 0xFFu, g = (c0 >> 8) & 0xFFu, b = (c0 >> 16) & 0xFFu, a = c0 >> 24;
            *o = make_uint4(0xFFFFFDFCu, 0xFFFFFFFFu, (r * 257u) | ((g * 257u) << 16), (b * 257u) | ((a * 257u) << 16));
            return;
        }
    }
    const int raw = sb.raw_mode & 0x7FF;
    const int wc = T.raw_wc[raw], q = T.raw_quant[raw];
    const int pc = sb.pc;
    // a block without any candidate (cannot happen for valid footprints) is
    // emitted as an all-zero block (reserved mode, decodes to the error colour)
    bool valid = sb.error < 1e30f && sb.raw_mode < 2048 && wc > 0 && pc >= 1 && pc <= 4 && sb.quant >= 4 &&
                 sb.quant <= 20;
    if (valid) {
        int nv = 0;
        for (int i = 0; i < pc; ++i) nv += 2 * (fmt_of(sb.formats, i) >> 2) + 2;
        const int n = kLevelsD[sb.quant];
        const int t = n % 3 == 0, u = n % 5 == 0;
        const int cb = 31 - __clz(t ? n / 3 : (u ? n / 5 : n));
        const int cbits = cb * nv + (t ? (8 * nv + 4) / 5 : 0) + (u ? (7 * nv + 2) / 3 : 0);
        const int wq = kLevelsD[q];
        const int wt = wq % 3 == 0, wu = wq % 5 == 0;
        const int wb = 31 - __clz(wt ? wq / 3 : (wu ? wq / 5 : wq));
        const int wbits_chk = wb * wc + (wt ? (8 * wc + 4) / 5 : 0) + (wu ? (7 * wc + 2) / 3 : 0);
        valid = nv <= 32 && (pc == 1 ? 17 : 29) + cbits + ((pc > 1 && !sb.matched) ? 3 * pc - 4 : 0) <= 128 - wbits_chk;
    }
    if (!valid) {
        *o = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    uint32_t wb[4] = {0, 0, 0, 0};
    uint8_t wv[kMaxWeights];
    const float levels = (float)kLevelsD[q];
    for (int i = 0; i < wc; ++i) {
        const float qw = ((float)min((int)sb.qw[i], 64) / 64.0f) * (levels - 1.0f);
        wv[i] = c_qt.weight_scramble[q][(int)(qw + 0.5f)];
    }
    encode_ise(q, wc, wv, wb, 0);
    // bit-reverse the whole 128-bit weight string into the top of the block
    uint32_t w[4];
    for (int i = 0; i < 4; ++i) w[i] = bitrev32(wb[3 - i]);
    int wbits;
    {
        const int n = kLevelsD[q];
        const int t = n % 3 == 0, u = n % 5 == 0;
        const int base = t ? n / 3 : (u ? n / 5 : n);
        const int b = 31 - __clz(base);
        wbits = b * wc + (t ? (8 * wc + 4) / 5 : 0) + (u ? (7 * wc + 2) / 3 : 0);
    }
    write_bits((uint32_t)raw, 11, 0, w);
    write_bits((uint32_t)(pc - 1), 2, 11, w);
    if (pc > 1) {
        write_bits(sb.seed & 0x3F, 6, 13, w);
        write_bits(sb.seed >> 6, 4, 19, w);
        if (sb.matched) {
            write_bits((uint32_t)fmt_of(sb.formats, 0) << 2, 6, 23, w);
        } else {
            int low_class = 4;
            for (int i = 0; i < pc; ++i) low_class = min(low_class, fmt_of(sb.formats, i) >> 2);
            if (low_class == 3) low_class = 2;
            int enc = low_class + 1, bitpos = 2;
            for (int i = 0; i < pc; ++i) { enc |= ((fmt_of(sb.formats, i) >> 2) - low_class) << bitpos; ++bitpos; }
            for (int i = 0; i < pc; ++i) { enc |= (fmt_of(sb.formats, i) & 3) << bitpos; bitpos += 2; }
            const int hsize = 3 * pc - 4;
            write_bits((uint32_t)(enc & 0x3F), 6, 23, w);
            write_bits((uint32_t)(enc >> 6), hsize, 128 - wbits - hsize, w);
        }
    } else {
        write_bits((uint32_t)fmt_of(sb.formats, 0), 4, 13, w);
        // dual plane (optional search): the plane-2 channel just below the weights
        if (sb.pad) write_bits((uint32_t)(sb.pad - 1) & 3u, 2, 128 - wbits - 2, w);
    }
    uint8_t vals[32];
    int nv = 0;
    const int cq = sb.quant;
    for (int i = 0; i < pc; ++i) {
        const int n = 2 * (fmt_of(sb.formats, i) >> 2) + 2;
        for (int j = 0; j < n; ++j) vals[nv + j] = c_qt.color_pquant[cq - 4][sb.packed[i * 8 + j]];
        nv += n;
    }
    encode_ise(cq, nv, vals, w, pc == 1 ? 17 : 29);
    *o = make_uint4(w[0], w[1], w[2], w[3]);
}


// ---- optional dual-plane search (ASTCG_FLAG_DUAL_PLANE) ----------------------
// The reference never tries dual-plane block modes (compression_pipeline.cpp:
// 183-185).  One lane per block, after the search: for every plane-2 channel
// c and every dual-plane mode of the footprint (single partition, direct
// endpoints: CEM 8, or CEM 12 when alpha varies), plane-1 weights come from
// a line fit of the other channels and plane-2 weights from channel c's own
// range; both are decimated to the mode's grid and quantised, the endpoints
// are refitted by least squares to the quantised weights and quantised to the
// colour level the mode leaves room for, and the candidate is scored with the
// search's final-error metric (pass17).  A candidate below the block's best
// error replaces it; SymBest::pad then holds c + 1.
constexpr int kDualMaxTc = 64;

size_t dual_lds_bytes(const DevTables& T) { return (size_t)2 * T.tc * 64 * sizeof(float) + 2 * 32 * 64; }

// Weight of texel t as the ASTC decoder infills it: (sum of grid weight x
// contribution in sixteenths + 8) >> 4 (contributions are multiples of 1/16);
// mirror: of the grid weights 64 - q (the rounding does not commute with
// the mirroring, so it is applied before the infill, as the packed block
// holds it).
__device__ __forceinline__ int infill_int(const DevTables& T, int dm, const uint8_t* qw, int stride, int t,
                                          bool mirror = false) {
    const int tc = T.tc;
    if (T.dm[dm].weight_count == tc) return mirror ? 64 - qw[t * stride] : qw[t * stride];
    const int n = T.t2w_count[(size_t)dm * tc + t];
    int s = 0;
    for (int j = 0; j < n; ++j) {
        const int q = qw[T.t2w_idx[((size_t)dm * tc + t) * 4 + j] * stride];
        s += (mirror ? 64 - q : q) * (int)(T.t2w_w[((size_t)dm * tc + t) * 4 + j] * 16.0f + 0.5f);
    }
    return (s + 8) >> 4;
}

// 8-bit value of a decoded texel channel: the 16-bit interpolation of the
// x257 endpoints, rounded to 8 bits.
__device__ __forceinline__ int decoded_u8(int e0, int e1, int w) {
    const int v16 = (e0 * 257 * (64 - w) + e1 * 257 * w + 32) >> 6;
    return (v16 * 255 + 32767) / 65535;
}

// Channel weights of the error metric (astc.h:60-63: 0.675, 1.3275, 0.2475,
// 1) times 4000, so that decoded 8-bit errors compare exactly in integers.
__device__ __constant__ uint32_t kCW4000[4] = {2700u, 5310u, 990u, 4000u};

__device__ __forceinline__ uint64_t texel_err_u8(const int* dec, uint32_t v) {
    uint64_t e = 0;
    for (int j = 0; j < 4; ++j) {
        const int d = dec[j] - (int)((v >> (8 * j)) & 0xFFu);
        e += (uint64_t)kCW4000[j] * (uint64_t)(d * d);
    }
    return e;
}

// Channel-weighted squared error (x4000, exact), in 8-bit units, of a
// single-plane symbolic block as the decoder reconstructs it (integer weight
// infill, decoded endpoints, 8-bit rounding) against the block's texels.
__device__ uint64_t decoded_block_error(const BlockLoc& loc, const DevTables& T, const SymBest& sb) {
    const int tc = T.tc, pc = sb.pc;
    if (pc < 1 || sb.raw_mode >= 2048 || T.raw_to_bm[sb.raw_mode] < 0) return ~0ull;
    const int dm = T.bm[T.raw_to_bm[sb.raw_mode]].dm;
    int e0[4][4], e1[4][4];
    for (int p = 0; p < pc; ++p) {
        uint32_t v[8];
        for (int i = 0; i < 8; ++i) v[i] = sb.packed[p * 8 + i];
        unpack_color_endpoints(fmt_of(sb.formats, p), v, e0[p], e1[p]);
    }
    const uint8_t* pot = nullptr;
    if (pc > 1) {
        const int e = T.seed_to_packed[(pc - 2) * kPartitionings + sb.seed];
        pot = T.part_of_texel + ((size_t)(pc - 2) * kPartitionings + e) * tc;
    }
    uint64_t total = 0;
    for (int t = 0; t < tc; ++t) {
        const int p = pot ? pot[t] : 0, w = infill_int(T, dm, sb.qw, 1, t);
        int dec[4];
        for (int j = 0; j < 4; ++j) dec[j] = decoded_u8(e0[p][j], e1[p][j], w);
        total += texel_err_u8(dec, texel_rgba8(loc, T, t));
    }
    return total;
}

__global__ void __launch_bounds__(64) k_dual(ImageArgs A, DevTables T, SymBest* best) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = threadIdx.x, k = blockIdx.x * 64 + lane, tc = T.tc;
    float* s_w = reinterpret_cast<float*>(smem);                  // [2][tc][64] ideal weights
    uint8_t* s_q = reinterpret_cast<uint8_t*>(s_w + 2 * tc * 64);  // [2][32][64] quantised grid weights
    if (k >= A.nb) return;  // no barriers below
    const int g = A.blk0 + k;
    const BlockLoc loc = block_loc(A, T, g);
    SymBest& sb = best[g];
    // candidates compete with the searched block in what the decoder
    // reconstructs (8-bit texels), so a replaced block never decodes worse
    uint64_t best_err = decoded_block_error(loc, T, sb);
    float* w1s = s_w + lane;
    float* w2s = s_w + tc * 64 + lane;
    uint8_t* q1s = s_q + lane;
    uint8_t* q2s = s_q + 32 * 64 + lane;
    bool alpha_opaque = true;
    for (int t = 0; t < tc; ++t) alpha_opaque &= (texel_rgba8(loc, T, t) >> 24) == 255u;
    const int cem = alpha_opaque ? 8 : 12, nv = alpha_opaque ? 6 : 8;
    for (int c = 0; c < 4; ++c) {
        // plane 2: channel c over its range
        float lo = 255.0f, hi = 0.0f;
        for (int t = 0; t < tc; ++t) {
            const float v = (float)((texel_rgba8(loc, T, t) >> (8 * c)) & 0xFFu);
            lo = fminf(lo, v);
            hi = fmaxf(hi, v);
        }
        if (hi <= lo) continue;  // a constant channel needs no plane of its own
        // plane 1: principal direction of the other channels (power iteration)
        float m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int t = 0; t < tc; ++t) {
            const uint32_t v = texel_rgba8(loc, T, t);
            for (int j = 0; j < 4; ++j) m[j] += (float)((v >> (8 * j)) & 0xFFu);
        }
        for (int j = 0; j < 4; ++j) m[j] = j == c ? 0.0f : m[j] / (float)tc;
        float cov[4][4] = {};
        for (int t = 0; t < tc; ++t) {
            const uint32_t v = texel_rgba8(loc, T, t);
            float d[4];
            for (int j = 0; j < 4; ++j) d[j] = j == c ? 0.0f : (float)((v >> (8 * j)) & 0xFFu) - m[j];
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) cov[i][j] += d[i] * d[j];
        }
        float dir[4] = {1.0f, 1.0f, 1.0f, 1.0f};
        dir[c] = 0.0f;
        for (int it = 0; it < 6; ++it) {
            float nd[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) nd[i] += cov[i][j] * dir[j];
            const float n2 = nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2] + nd[3] * nd[3];
            if (!(n2 > 1e-20f)) break;
            const float r = rsqrtf(n2);
            for (int i = 0; i < 4; ++i) dir[i] = nd[i] * r;
        }
        float smin = 3.4e38f, smax = -3.4e38f;
        for (int t = 0; t < tc; ++t) {
            const uint32_t v = texel_rgba8(loc, T, t);
            float sp = 0.0f;
            for (int j = 0; j < 4; ++j)
                if (j != c) sp += ((float)((v >> (8 * j)) & 0xFFu) - m[j]) * dir[j];
            w1s[t * 64] = sp;
            smin = fminf(smin, sp);
            smax = fmaxf(smax, sp);
        }
        const float span = smax - smin, inv1 = span > 1e-6f ? 1.0f / span : 0.0f, inv2 = 1.0f / (hi - lo);
        for (int t = 0; t < tc; ++t) {
            w1s[t * 64] = (w1s[t * 64] - smin) * inv1;
            w2s[t * 64] = ((float)((texel_rgba8(loc, T, t) >> (8 * c)) & 0xFFu) - lo) * inv2;
        }
        for (int di = 0; di < T.ndual; ++di) {
            const DualMode md = T.dual[di];
            const int cq = alpha_opaque ? md.cq_rgb : md.cq_rgba;
            if (cq < 4) continue;
            const DmInfo dinfo = T.dm[md.dm];
            const int wc = dinfo.weight_count, q = md.quant, lv = kLevelsD[q];
            const bool undecimated = wc == tc;
            const uint8_t* t2wc = T.t2w_count + (size_t)md.dm * tc;
            const uint8_t* t2wi = T.t2w_idx + (size_t)md.dm * tc * 4;
            const float* t2ww = T.t2w_w + (size_t)md.dm * tc * 4;
            const uint8_t* w2tc = T.w2t_count + md.dm * 64;
            const uint16_t* w2to = T.w2t_off + md.dm * 64;
            const uint8_t* w2ti = T.w2t_idx + dinfo.w2t_base;
            const float* w2tw = T.w2t_w + dinfo.w2t_base;
            // decimate and quantise both planes
            for (int plane = 0; plane < 2; ++plane) {
                const float* ws = plane ? w2s : w1s;
                uint8_t* qs = plane ? q2s : q1s;
                for (int w = 0; w < wc; ++w) {
                    float v;
                    if (undecimated) {
                        v = ws[w * 64];
                    } else {
                        float acc = 0.0f, den = 1e-10f;
                        const int n = w2tc[w], off = w2to[w];
                        for (int j = 0; j < n; ++j) {
                            acc += ws[w2ti[off + j] * 64] * w2tw[off + j];
                            den += w2tw[off + j];
                        }
                        v = acc / den;
                    }
                    const int rank = min(max((int)(fminf(fmaxf(v, 0.0f), 1.0f) * (float)(lv - 1) + 0.5f), 0), lv - 1);
                    qs[w * 64] = c_qt.quant_to_unquant[q][rank];
                }
            }
            // infilled weight (0..64) of texel t on a plane, as the decoder forms it
            auto infill = [&](const uint8_t* qs, int t, bool mirror = false) -> int {
                return infill_int(T, md.dm, qs, 64, t, mirror);
            };
            (void)t2wc;
            (void)t2wi;
            (void)t2ww;
            // least-squares endpoints per channel for the quantised weights
            float a00[2] = {0.0f, 0.0f}, a01[2] = {0.0f, 0.0f}, a11[2] = {0.0f, 0.0f};
            float b0[4] = {0.0f, 0.0f, 0.0f, 0.0f}, b1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int t = 0; t < tc; ++t) {
                const float wp[2] = {(float)infill(q1s, t) / 64.0f, (float)infill(q2s, t) / 64.0f};
                for (int p = 0; p < 2; ++p) {
                    a00[p] += (1.0f - wp[p]) * (1.0f - wp[p]);
                    a01[p] += (1.0f - wp[p]) * wp[p];
                    a11[p] += wp[p] * wp[p];
                }
                const uint32_t v = texel_rgba8(loc, T, t);
                for (int j = 0; j < 4; ++j) {
                    const float x = (float)((v >> (8 * j)) & 0xFFu), wj = wp[j == c ? 1 : 0];
                    b0[j] += (1.0f - wj) * x;
                    b1[j] += wj * x;
                }
            }
            int eq0[4], eq1[4];
            for (int j = 0; j < 4; ++j) {
                const int p = j == c ? 1 : 0;
                float e0 = j == c ? lo : m[j] + dir[j] * smin, e1 = j == c ? hi : m[j] + dir[j] * smax;
                const float det = a00[p] * a11[p] - a01[p] * a01[p];
                if (fabsf(det) > 1e-4f * (a00[p] * a00[p] + a11[p] * a11[p] + 2.0f * a01[p] * a01[p])) {
                    e0 = (a11[p] * b0[j] - a01[p] * b1[j]) / det;
                    e1 = (a00[p] * b1[j] - a01[p] * b0[j]) / det;
                }
                if (j == 3 && alpha_opaque) { e0 = 255.0f; e1 = 255.0f; }
                const int i0 = min(max((int)rintf(2.0f * e0), 0), 510), i1 = min(max((int)rintf(2.0f * e1), 0), 510);
                eq0[j] = c_qt.color_u2q[cq - 4][i0];
                eq1[j] = c_qt.color_u2q[cq - 4][i1];
            }
            // direct endpoints decode with blue contraction when the second
            // endpoint's RGB sum is the smaller: swap the endpoints and
            // mirror both planes' weights instead (weight levels are
            // symmetric about 32)
            const bool swap = eq1[0] + eq1[1] + eq1[2] < eq0[0] + eq0[1] + eq0[2];
            if (swap)
                for (int j = 0; j < 4; ++j) { const int x = eq0[j]; eq0[j] = eq1[j]; eq1[j] = x; }
            // decoded error (8-bit units, exact integers) with an early
            // exit, and the search's error (pass17 scale) of the same
            // reconstruction
            uint64_t total = 0;
            float total16 = 0.0f;
            for (int t = 0; t < tc && total < best_err; ++t) {
                const int wt[2] = {infill(q1s, t, swap), infill(q2s, t, swap)};
                const uint32_t v = texel_rgba8(loc, T, t);
                float dd16[4];
                int dec[4];
                for (int j = 0; j < 4; ++j) {
                    const int wv = wt[j == c ? 1 : 0];
                    const uint32_t x = (v >> (8 * j)) & 0xFFu;
                    const int col = (eq0[j] * 257 * (64 - wv) + eq1[j] * 257 * wv + 32) >> 6;
                    const float df16 = wmin(fabsf(u8_search_value(x) - (float)col), 1e15f);
                    dd16[j] = df16 * df16;
                    dec[j] = (col * 255 + 32767) / 65535;
                }
                total += texel_err_u8(dec, v);
                total16 += wmin(dot4(dd16, kCW), kErrDefault);
            }
            if (!(total < best_err)) continue;
            best_err = total;
            sb.error = total16;
            sb.raw_mode = md.raw_mode;
            sb.pc = 1;
            sb.seed = 0;
            sb.matched = 0;
            sb.quant = (uint8_t)cq;
            sb.formats = (uint16_t)cem;
            sb.pad = (uint8_t)(c + 1);
            // CEM 8 / 12 value order: R0 R1 G0 G1 B0 B1 (A0 A1)
            for (int j = 0; j < nv / 2; ++j) {
                sb.packed[2 * j] = (uint8_t)eq0[j];
                sb.packed[2 * j + 1] = (uint8_t)eq1[j];
            }
            // weights interleaved plane 1, plane 2 per grid point
            for (int w = 0; w < wc; ++w) {
                const int u1 = q1s[w * 64], u2 = q2s[w * 64];
                sb.qw[2 * w] = (uint8_t)(swap ? 64 - u1 : u1);
                sb.qw[2 * w + 1] = (uint8_t)(swap ? 64 - u2 : u2);
            }
        }
    }
}

// ---- learned candidate policy: per-chunk search plan ------------------------
// Predictor outputs per block (top-k of its five heads): [nb][5][kmax]
// indices and probabilities; heads are partition count (1..4 as 0..3),
// seeds for 2 / 3 / 4 partitions, trial mode.
__global__ void __launch_bounds__(64) k_plan_zero(int* counts) {
    if (threadIdx.x < 3) counts[threadIdx.x] = 0;
}

// Zeroes n ints (a kernel rather than a memset, so that it can sit in a
// captured graph).
__global__ void __launch_bounds__(256) k_zero_ints(int* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void __launch_bounds__(64) k_fb_reset(unsigned long long* fb_count) {
    if (threadIdx.x == 0) *fb_count = 0ull;
}

// The weighted squared distance of a block's texels to their best line (the
// average plus the direction of the largest positive-deviation sum, as the
// partition search's line fits choose it), over RGBA when alpha varies, RGB
// otherwise: what one partition cannot encode whatever its endpoints.  A
// decision heuristic of the learned policy, no counterpart in the reference.
__device__ float block_line_error(const ImageArgs& A, const DevTables& T, int g) {
    const BlockLoc loc = block_loc(A, T, g);
    const int tc = T.tc;
    float avg[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t amin = 0xFFu, amax = 0u;
    for (int t = 0; t < tc; ++t) {
        float px[4];
        texel_px(loc, T, t, px);
        for (int c = 0; c < 4; ++c) avg[c] += px[c];
        const uint32_t a = texel_rgba8(loc, T, t) >> 24;
        amin = min(amin, a);
        amax = max(amax, a);
    }
    for (int c = 0; c < 4; ++c) avg[c] *= 1.0f / (float)tc;
    const int nch = amin != amax ? 4 : 3;
    float sx[4] = {}, sy[4] = {}, sz[4] = {}, sw[4] = {};
    for (int t = 0; t < tc; ++t) {
        float px[4], d[4];
        texel_px(loc, T, t, px);
        for (int c = 0; c < 4; ++c) d[c] = px[c] - avg[c];
        if (d[0] > 0) for (int c = 0; c < 4; ++c) sx[c] += d[c];
        if (d[1] > 0) for (int c = 0; c < 4; ++c) sy[c] += d[c];
        if (d[2] > 0) for (int c = 0; c < 4; ++c) sz[c] += d[c];
        if (nch == 4 && d[3] > 0) for (int c = 0; c < 4; ++c) sw[c] += d[c];
    }
    float dir[4], best = -1.0f;
    auto pick = [&](const float* v) {
        float n = 0.0f;
        for (int c = 0; c < nch; ++c) n += v[c] * v[c];
        if (n > best) {
            best = n;
            for (int c = 0; c < 4; ++c) dir[c] = c < nch ? v[c] : 0.0f;
        }
    };
    pick(sx);
    pick(sy);
    pick(sz);
    if (nch == 4) pick(sw);
    const float inv = best > 0.0f ? 1.0f / sqrtf(best) : 0.0f;
    for (int c = 0; c < 4; ++c) dir[c] *= inv;
    float err = 0.0f;
    for (int t = 0; t < tc; ++t) {
        float px[4], d[4];
        texel_px(loc, T, t, px);
        float u = 0.0f;
        for (int c = 0; c < nch; ++c) {
            d[c] = px[c] - avg[c];
            u += d[c] * dir[c];
        }
        for (int c = 0; c < nch; ++c) {
            const float r = d[c] - u * dir[c];
            err += kCW[c] * r * r;
        }
    }
    return err;
}

// One lane per block for the decisions; the blocks' mode-head entries are
// first staged through LDS 16 at a time with contiguous loads (a block's
// top-k record is 5 * kmax entries, so lane-per-block loads would touch one
// cache line per lane).  The kept probability is still summed per block in
// rank order.
__global__ void __launch_bounds__(64) k_plan_modes(ImageArgs A, DevTables T, SearchPlan P, int nbm, uint32_t* mask,
                                                   uint8_t* slots) {
    __shared__ int s_idx[64 * 17];    // [block][16 entries], rows padded to 17 words
    __shared__ float s_prob[64 * 17];
    const int base = blockIdx.x * 64, lane = threadIdx.x, k = base + lane;
    const bool live = k < A.nb;
    const int nblk = min(64, A.nb - base);
    // confidence: the probability mass of the trial modes the block keeps
    // (the top-k kernel fills min(modes, nbm) entries of the mode head)
    const int nkeep = min(min(P.kmax, P.modes), nbm);
    auto stage = [&](int j0, int nj, bool probs) {
        __syncthreads();
        for (int e = lane; e < 64 * 16; e += 64) {
            const int b = e >> 4, j = e & 15;
            if (b < nblk && j < nj) {
                const size_t o = (size_t)(base + b) * 5 * P.kmax + 4 * P.kmax + j0 + j;
                s_idx[b * 17 + j] = P.topk_idx[o];
                if (probs) s_prob[b * 17 + j] = P.topk_prob[o];
            }
        }
        __syncthreads();
    };
    float kept = 0.0f;
    for (int j0 = 0; j0 < nkeep; j0 += 16) {
        const int nj = min(16, nkeep - j0);
        stage(j0, nj, true);
        if (live)
            for (int j = 0; j < nj; ++j)
                if (s_idx[lane * 17 + j] >= 0) kept += s_prob[lane * 17 + j];
    }
    // blocks the pruning should not touch: low confidence, or (line_limit)
    // more error off their best line than the error limit allows
    const bool fb = live && (kept < P.fallback_below ||
                             (P.line_limit > 0.0f && block_line_error(A, T, A.blk0 + k) > P.line_limit));
    // one atomic per wave for the fallback statistics
    const unsigned long long wave_fb = __ballot(fb);
    if (lane == 0 && wave_fb) atomicAdd(P.fb_count, (unsigned long long)__popcll(wave_fb));
    uint32_t* m = mask + (size_t)k * P.mask_words;
    if (live) {
        P.fallback[k] = fb ? 1 : 0;
        for (int w = 0; w < P.mask_words; ++w) {
            const int lo = 32 * w;
            m[w] = !fb ? 0u : (nbm - lo >= 32 ? 0xFFFFFFFFu : (nbm > lo ? (1u << (nbm - lo)) - 1u : 0u));
        }
        slots[k] = fb ? 8 : (uint8_t)P.nslots;
    }
    if (__ballot(live && !fb) == 0ull) return;  // every block of the wave falls back
    for (int j0 = 0; j0 < nkeep; j0 += 16) {
        const int nj = min(16, nkeep - j0);
        stage(j0, nj, false);
        if (live && !fb)
            for (int j = 0; j < nj; ++j) {
                const int md = s_idx[lane * 17 + j];
                if (md >= 0 && md < nbm) m[md >> 5] |= 1u << (md & 31);
            }
    }
}

__device__ __forceinline__ float bf16_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t v) { return __uint_as_float(v & 0xFFFF0000u); }

__device__ __forceinline__ float dot8_bf16(uint4 a, uint4 b) {
    float r = bf16_lo(a.x) * bf16_lo(b.x);
    r += bf16_hi(a.x) * bf16_hi(b.x);
    r += bf16_lo(a.y) * bf16_lo(b.y);
    r += bf16_hi(a.y) * bf16_hi(b.y);
    r += bf16_lo(a.z) * bf16_lo(b.z);
    r += bf16_hi(a.z) * bf16_hi(b.z);
    r += bf16_lo(a.w) * bf16_lo(b.w);
    r += bf16_hi(a.w) * bf16_hi(b.w);
    return r;
}

// Keeps the P.keep_pb of block k's 4 compat positions (pb_part[k * 4 + c],
// packed table entries of partition count pc) whose seeds score highest in
// the predictor's seed head (logit = b3 + h2 . W3 column, FP32 sums of BF16
// products in index order; ties keep the lower position), compacted in
// position order into slots 0..n-1; returns n.
__device__ int rank_positions(const DevTables& T, const SearchPlan& P, int pc, int k, int16_t* pb_part) {
    const int H = P.hidden;
    int row[4];
    float acc[4];
    for (int c = 0; c < 4; ++c) {
        const int e = pb_part[k * 4 + c];
        row[c] = (e >= 0 && e < kPartitionings) ? (pc - 2) * kPartitionings + T.part_seed[(pc - 2) * kPartitionings + e]
                                                : -1;
        acc[c] = row[c] >= 0 ? P.b3_seed[row[c]] : -INFINITY;
    }
    const uint4* h = reinterpret_cast<const uint4*>(P.h2 + (size_t)k * H);
    for (int j = 0; j < H / 8; ++j) {
        const uint4 hv = h[j];
        for (int c = 0; c < 4; ++c)
            if (row[c] >= 0) acc[c] += dot8_bf16(hv, reinterpret_cast<const uint4*>(P.w3_seed + (size_t)row[c] * H)[j]);
    }
    int n = 0;
    for (int c = 0; c < 4; ++c) {
        int r = 0;
        for (int d = 0; d < 4; ++d) r += (acc[d] > acc[c]) || (acc[d] == acc[c] && d < c);
        if (r < P.keep_pb) pb_part[k * 4 + n++] = pb_part[k * 4 + c];
    }
    return n;
}

// One lane per block: the partitioned blocks of partition count pc.  A
// fallback block keeps the 4 compat positions k_partition wrote; another
// block searches, when the count is its most likely one or at least
// count_below probable, the compat positions (all 4, or the keep_pb its
// seed head ranks highest) or the predictor's top seeds of this count,
// written into its pb_part slots 0..n-1.  Entries are appended to the list
// through one atomic per workgroup; their order does not affect any result.
__global__ void __launch_bounds__(64) k_plan(ImageArgs A, DevTables T, int pc, int16_t* pb_part, SearchPlan P,
                                             int* count) {
    __shared__ int s_n, s_base;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    int n = 0;
    if (k < A.nb) {
        const int* ti = P.topk_idx + (size_t)k * 5 * P.kmax;
        const int g = A.blk0 + k;
        if (pc == 2 && !P.fallback[k] && P.part_slots != P.nslots)
            const_cast<uint8_t*>(P.slots)[k] = (uint8_t)P.part_slots;  // slots for 2..4 partitions
        if (plan_early_out(P, g)) {
            n = 0;  // early out: the block's best encoding so far is good enough
        } else if (plan_hard(P, pc, k, g)) {
            // a hard block: the full search from 2 partitions on (every
            // count, the full partition candidates, and unless
            // hard_keeps_modes all trial modes and 8 slots; later counts see
            // the fallback flag)
            P.fallback[k] = 1;
            atomicAdd(P.fb_count, 1ull);
            if (!P.hard_keeps_modes) {
                uint32_t* m = const_cast<uint32_t*>(P.mode_mask) + (size_t)k * P.mask_words;
                for (int w = 0; w < P.mask_words; ++w) {
                    const int lo = 32 * w;
                    m[w] = T.nbm - lo >= 32 ? 0xFFFFFFFFu : (T.nbm > lo ? (1u << (T.nbm - lo)) - 1u : 0u);
                }
                const_cast<uint8_t*>(P.slots)[k] = 8;
            }
            n = 4;
        } else if (P.fallback[k]) {
            n = 4;
        } else if (plan_keeps_count(P, pc, k, g)) {
            if (P.seeds == 0) {
                // the compat positions k_partition wrote
                n = P.keep_pb < 4 ? rank_positions(T, P, pc, k, pb_part) : 4;
            } else if (P.seeds > 4) {
                // the best of its seeds by line fit (k_partition_ex), -1 after the last
                while (n < 4 && pb_part[k * 4 + n] >= 0) ++n;
            } else {
                for (int j = 0; j < P.kmax && j < P.seeds && n < 4; ++j) {
                    const int seed = ti[(pc - 1) * P.kmax + j];
                    if (seed < 0 || seed >= kPartitionings) break;
                    const int e = T.seed_to_packed[(pc - 2) * kPartitionings + seed];
                    if (e >= T.count_selected[pc - 1]) continue;  // not a selected partitioning
                    pb_part[k * 4 + n] = (int16_t)e;
                    ++n;
                }
            }
        }
    }
    const int local = n ? atomicAdd(&s_n, n) : 0;
    __syncthreads();
    if (threadIdx.x == 0) s_base = s_n ? atomicAdd(count, s_n) : 0;
    __syncthreads();
    if (k < A.nb) {
        for (int c = 0; c < 4; ++c) P.pb_inv[k * 4 + c] = -1;
        for (int c = 0; c < n; ++c) {
            const int idx = s_base + local + c;
            P.pb_list[idx] = k * 4 + c;
            P.pb_inv[k * 4 + c] = idx;
        }
    }
}

// ---- ASTC LDR decoder (Khronos ASTC specification, 2D LDR profile) ---------
// Independent of the search: one lane per block, writing the decoded RGBA8
// texels that fall inside the image.  Error blocks and HDR endpoint modes
// decode to the error colour (magenta), LDR void-extent blocks to their
// constant colour.
__device__ __forceinline__ uint32_t rd_bits(const uint32_t* w, int off, int n) {
    if (n <= 0) return 0u;
    const int wi = off >> 5, sh = off & 31;
    uint64_t v = w[wi];
    if (wi < 3) v |= (uint64_t)w[wi + 1] << 32;
    return (uint32_t)(v >> sh) & (uint32_t)((1ull << n) - 1ull);
}

__device__ __forceinline__ void ise_shape(int q, int* bits, int* trits, int* quints) {
    const int n = kLevelsD[q];
    *trits = n % 3 == 0;
    *quints = n % 5 == 0;
    const int base = *trits ? n / 3 : (*quints ? n / 5 : n);
    *bits = 31 - __clz(base);
}

__device__ int ise_bits(int count, int q) {
    int b, t, u;
    ise_shape(q, &b, &t, &u);
    return b * count + (t ? (8 * count + 4) / 5 : 0) + (u ? (7 * count + 2) / 3 : 0);
}

__device__ void ise_decode(int q, int count, const uint32_t* w, int off, uint8_t* out) {
    int b, t, u;
    ise_shape(q, &b, &t, &u);
    if (t) {
        const int tb[5] = {2, 2, 1, 2, 1}, ts[5] = {0, 2, 4, 5, 7};
        for (int i = 0; i < count; i += 5) {
            int m[5] = {0, 0, 0, 0, 0}, T = 0;
            for (int k = 0; k < 5 && i + k < count; ++k) {
                m[k] = (int)rd_bits(w, off, b);
                off += b;
                T |= (int)rd_bits(w, off, tb[k]) << ts[k];
                off += tb[k];
            }
            const int dg = c_qt.trit_digits[T];
            for (int k = 0; k < 5 && i + k < count; ++k) out[i + k] = (uint8_t)((((dg >> (2 * k)) & 3) << b) | m[k]);
        }
    } else if (u) {
        const int qb[3] = {3, 2, 2}, qs[3] = {0, 3, 5};
        for (int i = 0; i < count; i += 3) {
            int m[3] = {0, 0, 0}, Q = 0;
            for (int k = 0; k < 3 && i + k < count; ++k) {
                m[k] = (int)rd_bits(w, off, b);
                off += b;
                Q |= (int)rd_bits(w, off, qb[k]) << qs[k];
                off += qb[k];
            }
            const int dg = c_qt.quint_digits[Q];
            for (int k = 0; k < 3 && i + k < count; ++k) out[i + k] = (uint8_t)((((dg >> (3 * k)) & 7) << b) | m[k]);
        }
    } else {
        for (int i = 0; i < count; ++i) {
            out[i] = (uint8_t)rd_bits(w, off, b);
            off += b;
        }
    }
}

// Block-mode field decode (ASTC spec, 2D): weight grid, dual plane, weight quant.
__device__ bool dec_mode_fields(int mode, int* xw, int* yw, int* dual, int* quant) {
    int base_q = (mode >> 4) & 1;
    int H = (mode >> 9) & 1, D = (mode >> 10) & 1;
    const int A = (mode >> 5) & 3;
    if (mode & 3) {
        base_q |= (mode & 3) << 1;
        int B = (mode >> 7) & 3;
        switch ((mode >> 2) & 3) {
            case 0: *xw = B + 4; *yw = A + 2; break;
            case 1: *xw = B + 8; *yw = A + 2; break;
            case 2: *xw = A + 2; *yw = B + 8; break;
            default:
                B &= 1;
                if (mode & 0x100) { *xw = B + 2; *yw = A + 2; }
                else { *xw = A + 2; *yw = B + 6; }
                break;
        }
    } else {
        base_q |= ((mode >> 2) & 3) << 1;
        if (((mode >> 2) & 3) == 0) return false;
        const int B = (mode >> 9) & 3;
        switch ((mode >> 7) & 3) {
            case 0: *xw = 12; *yw = A + 2; break;
            case 1: *xw = A + 2; *yw = 12; break;
            case 2: *xw = A + 6; *yw = B + 6; D = 0; H = 0; break;
            default:
                if (((mode >> 5) & 3) == 0) { *xw = 6; *yw = 10; }
                else if (((mode >> 5) & 3) == 1) { *xw = 10; *yw = 6; }
                else return false;
                break;
        }
    }
    *quant = (base_q - 2) + 6 * H;
    *dual = D;
    return true;
}

__device__ __forceinline__ void bit_transfer_signed(int& a, int& b) {
    b >>= 1;
    b |= a & 0x80;
    a >>= 1;
    a &= 0x3F;
    if (a & 0x20) a -= 0x40;
}
__device__ __forceinline__ int clamp255(int x) { return x < 0 ? 0 : (x > 255 ? 255 : x); }

// LDR colour endpoint decode (endpoint modes 0, 1, 4, 5, 6, 8, 9, 10, 12, 13).
__device__ bool dec_endpoints(int cem, const int* v, int* e0, int* e1) {
    switch (cem) {
        case 0:
            e0[0] = e0[1] = e0[2] = v[0]; e0[3] = 255;
            e1[0] = e1[1] = e1[2] = v[1]; e1[3] = 255;
            return true;
        case 1: {
            const int l0 = (v[0] >> 2) | (v[1] & 0xC0);
            const int l1 = min(l0 + (v[1] & 0x3F), 255);
            e0[0] = e0[1] = e0[2] = l0; e0[3] = 255;
            e1[0] = e1[1] = e1[2] = l1; e1[3] = 255;
            return true;
        }
        case 4:
            e0[0] = e0[1] = e0[2] = v[0]; e0[3] = v[2];
            e1[0] = e1[1] = e1[2] = v[1]; e1[3] = v[3];
            return true;
        case 5: {
            int v0 = v[0], v1 = v[1], v2 = v[2], v3 = v[3];
            bit_transfer_signed(v1, v0);
            bit_transfer_signed(v3, v2);
            e0[0] = e0[1] = e0[2] = v0; e0[3] = v2;
            e1[0] = e1[1] = e1[2] = clamp255(v0 + v1); e1[3] = clamp255(v2 + v3);
            return true;
        }
        case 6:
        case 10:
            for (int c = 0; c < 3; ++c) { e1[c] = v[c]; e0[c] = (v[c] * v[3]) >> 8; }
            e0[3] = cem == 10 ? v[4] : 255;
            e1[3] = cem == 10 ? v[5] : 255;
            return true;
        case 8:
        case 12: {
            int c0[4] = {v[0], v[2], v[4], cem == 12 ? v[6] : 255};
            int c1[4] = {v[1], v[3], v[5], cem == 12 ? v[7] : 255};
            if (c1[0] + c1[1] + c1[2] < c0[0] + c0[1] + c0[2]) {  // blue contraction
                for (int c = 0; c < 2; ++c) { c0[c] = (c0[c] + c0[2]) >> 1; c1[c] = (c1[c] + c1[2]) >> 1; }
                for (int c = 0; c < 4; ++c) { e0[c] = c1[c]; e1[c] = c0[c]; }
            } else {
                for (int c = 0; c < 4; ++c) { e0[c] = c0[c]; e1[c] = c1[c]; }
            }
            return true;
        }
        case 9:
        case 13: {
            int a[4] = {v[0], v[2], v[4], cem == 13 ? v[6] : 255};
            int b[4] = {v[1], v[3], v[5], cem == 13 ? v[7] : 0};
            for (int c = 0; c < (cem == 13 ? 4 : 3); ++c) bit_transfer_signed(b[c], a[c]);
            int c0[4], c1[4];
            for (int c = 0; c < 4; ++c) { c0[c] = a[c]; c1[c] = a[c] + b[c]; }
            if (cem == 9) { c0[3] = 255; c1[3] = 255; }
            if (b[0] + b[1] + b[2] < 0) {
                for (int c = 0; c < 2; ++c) { c0[c] = (c0[c] + c0[2]) >> 1; c1[c] = (c1[c] + c1[2]) >> 1; }
                for (int c = 0; c < 4; ++c) { e0[c] = clamp255(c1[c]); e1[c] = clamp255(c0[c]); }
            } else {
                for (int c = 0; c < 4; ++c) { e0[c] = clamp255(c0[c]); e1[c] = clamp255(c1[c]); }
            }
            return true;
        }
        default:
            return false;  // HDR endpoint modes
    }
}

__global__ void __launch_bounds__(64) k_decode(DevTables T, const uint8_t* blocks, int width, int height, int blocks_x,
                                               int nblocks, uint8_t* rgba) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nblocks) return;
    const int bx = T.bx, by = T.by, tc = T.tc;
    const int x0 = (k % blocks_x) * bx, y0 = (k / blocks_x) * by;
    const uint4 raw = reinterpret_cast<const uint4*>(blocks)[k];
    const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
    auto put = [&](int t, uint32_t px) {
        const int x = x0 + t % bx, y = y0 + t / bx;
        if (x < width && y < height) *reinterpret_cast<uint32_t*>(rgba + ((size_t)y * width + x) * 4) = px;
    };
    const uint32_t magenta = 0xFFFF00FFu;
    auto fill = [&](uint32_t px) { for (int t = 0; t < tc; ++t) put(t, px); };
    const int mode = (int)rd_bits(w, 0, 11);
    if ((mode & 0x1FF) == 0x1FC) {
        if (mode & 0x200) { fill(magenta); return; }  // HDR void extent
        uint32_t px = 0;
        for (int c = 0; c < 4; ++c) px |= (uint32_t)((rd_bits(w, 64 + 16 * c, 16) + 128) / 257) << (8 * c);
        fill(px);
        return;
    }
    int xw = 0, yw = 0, dual = 0, q = 0;
    if (!dec_mode_fields(mode, &xw, &yw, &dual, &q) || xw > bx || yw > by) { fill(magenta); return; }
    const int wc = xw * yw * (dual + 1);
    const int wbits = ise_bits(wc, q);
    if (wc > 64 || wbits < 24 || wbits > 96) { fill(magenta); return; }
    const int pc = (int)rd_bits(w, 11, 2) + 1;
    if (dual && pc == 4) { fill(magenta); return; }
    int below = 128 - wbits, color_start, seed = 0;
    int cem[4] = {0, 0, 0, 0};
    if (pc == 1) {
        cem[0] = (int)rd_bits(w, 13, 4);
        color_start = 17;
    } else {
        seed = (int)rd_bits(w, 13, 10);
        color_start = 29;
        const int cemlow = (int)rd_bits(w, 23, 6);
        if ((cemlow & 3) == 0) {
            for (int i = 0; i < pc; ++i) cem[i] = cemlow >> 2;
        } else {
            const int hsize = 3 * pc - 4;
            below -= hsize;
            const int enc = cemlow | ((int)rd_bits(w, below, hsize) << 6);
            const int base_class = (enc & 3) - 1;
            int bitpos = 2, cls[4];
            for (int i = 0; i < pc; ++i) cls[i] = base_class + ((enc >> bitpos++) & 1);
            for (int i = 0; i < pc; ++i) { cem[i] = (cls[i] << 2) | ((enc >> bitpos) & 3); bitpos += 2; }
        }
    }
    int ccs = 0;
    if (dual) { below -= 2; ccs = (int)rd_bits(w, below, 2); }
    int nvals = 0;
    for (int i = 0; i < pc; ++i) nvals += 2 * ((cem[i] >> 2) + 1);
    if (nvals > 18) { fill(magenta); return; }
    const int avail = below - color_start;
    int cq = -1;
    for (int qq = 4; qq <= 20; ++qq)
        if (ise_bits(nvals, qq) <= avail) cq = qq;
    if (cq < 0) { fill(magenta); return; }
    uint8_t codes[18];
    ise_decode(cq, nvals, w, color_start, codes);
    int vals[18];
    for (int i = 0; i < nvals; ++i) vals[i] = c_qt.color_unq[cq - 4][codes[i]];
    int e0[4][4], e1[4][4];
    int vi = 0;
    for (int p = 0; p < pc; ++p) {
        if (!dec_endpoints(cem[p], vals + vi, e0[p], e1[p])) { fill(magenta); return; }
        vi += 2 * ((cem[p] >> 2) + 1);
    }
    // weights are stored bit-reversed from the top of the block
    const uint32_t rw[4] = {__brev(w[3]), __brev(w[2]), __brev(w[1]), __brev(w[0])};
    uint8_t wcodes[64];
    ise_decode(q, wc, rw, 0, wcodes);
    int uw[64];
    for (int i = 0; i < wc; ++i) uw[i] = c_qt.weight_unq[q][wcodes[i]];
    const int planes = dual ? 2 : 1;
    const uint8_t* part = nullptr;
    if (pc > 1) {
        const int e = T.seed_to_packed[(pc - 2) * kPartitionings + seed];
        part = T.part_of_texel + ((size_t)(pc - 2) * kPartitionings + e) * tc;
    }
    for (int t = 0; t < tc; ++t) {
        const int y = t / bx, x = t - y * bx;
        // weight infill
        const int xwp = (((1024 + bx / 2) / (bx - 1)) * x * (xw - 1) + 32) >> 6;
        const int ywp = (((1024 + by / 2) / (by - 1)) * y * (yw - 1) + 32) >> 6;
        const int xf = xwp & 0xF, yf = ywp & 0xF, base = (xwp >> 4) + (ywp >> 4) * xw;
        const int w11 = (xf * yf + 8) >> 4, w01 = xf - w11, w10 = yf - w11, w00 = 16 - xf - yf + w11;
        int wt[2] = {0, 0};
        for (int plane = 0; plane < planes; ++plane) {
            auto W = [&](int idx) { return idx < xw * yw ? uw[idx * planes + plane] : 0; };
            int s = 0;
            if (w00) s += W(base) * w00;
            if (w01) s += W(base + 1) * w01;
            if (w10) s += W(base + xw) * w10;
            if (w11) s += W(base + xw + 1) * w11;
            wt[plane] = (s + 8) >> 4;
        }
        const int p = part ? part[t] : 0;
        uint32_t px = 0;
        for (int c = 0; c < 4; ++c) {
            const int wv = (dual && c == ccs) ? wt[1] : wt[0];
            const int c0 = e0[p][c] * 257, c1 = e1[p][c] * 257;
            const int v16 = (c0 * (64 - wv) + c1 * wv + 32) >> 6;
            px |= (uint32_t)((v16 * 255 + 32767) / 65535) << (8 * c);
        }
        put(t, px);
    }
}

}  // namespace

const char* kernel_name(int id) {
    static const char* names[K_COUNT] = {"k_partition", "k_ideal_color", "k_dm_bm", "k_top",
                                         "k_refine", "k_pick", "k_pack"};
    return id >= 0 && id < K_COUNT ? names[id] : "?";
}

#define ASTCG_LAUNCH(id, kernel, grid, ...)                                        \
    do {                                                                           \
        if (prof) prof->begin(id, s);                                              \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), 0, s, __VA_ARGS__);  \
        if (prof) prof->end(id, s);                                                \
    } while (0)

// Workgroups of a persistent kernel: as many as the device holds at once
// (per kernel and LDS size, looked up once per device).
unsigned resident_workgroups(const void* kernel, size_t lds) {
    struct Entry { const void* k; size_t lds; int dev; unsigned n; };
    static std::mutex mu;
    static std::vector<Entry> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lock(mu);
    for (const Entry& e : cache)
        if (e.k == kernel && e.lds == lds && e.dev == dev) return e.n;
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1 ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, lds) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        return 1024;  // not cached: the queue is correct at any grid size
    }
    cache.push_back(Entry{kernel, lds, dev, (unsigned)(cus * per_cu)});
    return (unsigned)(cus * per_cu);
}

template <bool kStage>
void launch_refine_pc(int pc, unsigned tiles, size_t lds, hipStream_t s, const ImageArgs& A, const DevTables& T,
                      const ChunkBuffers& B, int npb, int iterations) {
    if (T.refine_queue > 0) {
        const int refill = T.refine_queue < 64 ? T.refine_queue : 64;
#define ASTCG_REFINE_Q(PCV)                                                                                        \
    do {                                                                                                          \
        const unsigned n = resident_workgroups(reinterpret_cast<const void*>(&k_refine_q<kStage, PCV>), lds);     \
        const unsigned grid = std::max(1u, std::min(tiles, n));                                                   \
        hipLaunchKernelGGL((k_refine_q<kStage, PCV>), dim3(grid), dim3(64), lds, s, A, T, B.pb_part, npb, B.ideal, \
                           B.top, B.ccr, B.qw, B.work, B.work_total, B.tile_next, B.cand, B.cand_err, B.wstride,  \
                           iterations, refill);                                                                  \
    } while (0)
        switch (pc) {
            case 1: ASTCG_REFINE_Q(1); break;
            case 2: ASTCG_REFINE_Q(2); break;
            case 3: ASTCG_REFINE_Q(3); break;
            default: ASTCG_REFINE_Q(4); break;
        }
#undef ASTCG_REFINE_Q
        return;
    }
#define ASTCG_REFINE(PCV)                                                                                         \
    hipLaunchKernelGGL((k_refine<kStage, PCV>), dim3(tiles), dim3(64), lds, s, A, T, B.pb_part, npb, B.ideal, B.top, \
                       B.ccr, B.qw, B.work, B.work_total, B.cand, B.cand_err, B.wstride, iterations)
    switch (pc) {
        case 1: ASTCG_REFINE(1); break;
        case 2: ASTCG_REFINE(2); break;
        case 3: ASTCG_REFINE(3); break;
        default: ASTCG_REFINE(4); break;
    }
#undef ASTCG_REFINE
}

// tiles: an upper bound of the work list's 64-entry tiles (the list length
// is only known on the device)
void launch_refine(bool stage, int pc, unsigned tiles, size_t lds, hipStream_t s, const ImageArgs& A,
                   const DevTables& T, const ChunkBuffers& B, int npb, int iterations) {
    if (stage) launch_refine_pc<true>(pc, tiles, lds, s, A, T, B, npb, iterations);
    else launch_refine_pc<false>(pc, tiles, lds, s, A, T, B, npb, iterations);
}

// ---- exhaustive partition candidates of a short block list ----------------
// The same search as k_partition_ex (bit for bit: every float operation in
// the same order), one wave per listed block instead of one lane: the lanes
// share the block's texels in LDS and split the texel-parallel steps (k-means
// assignment, the k-means texels' cluster masks) and the partitioning-parallel
// ones (mismatch counts, the counting sort's placement, the line fits); the
// sequential steps (k-means seeding and centre sums, the prefix of the
// histogram, the top-4 insertion in mismatch order) run on every lane alike.
// The workgroups loop over the list, so a list of a few hundred blocks (the
// learned policy's hard and fallback blocks) fills the chip.
constexpr int kExwGrid = 4096;

size_t partition_exw_lds_bytes(const DevTables& T) {
    return (size_t)T.tc * 16 + (((size_t)T.tc + 15) & ~size_t(15)) + kPartitionings * (1 + 4 + 4 + 4) + 2 * 64 * 4;
}

template <int PC, int kW>
__global__ void __launch_bounds__(64) k_partition_exw(ImageArgs A, DevTables T, int16_t* pb_part, const int* list,
                                                      const int* list_count, int search_limit) {
    constexpr int pc = PC;
    extern __shared__ __align__(16) uint8_t smem[];
    const int tc = T.tc, lane = threadIdx.x;
    float4* s_px = reinterpret_cast<float4*>(smem);                                    // [tc]
    uint8_t* s_part = smem + (size_t)tc * 16;                                          // [tc] k-means cluster
    uint8_t* s_mis = s_part + ((tc + 15) & ~15);                                       // [1024] mismatch / 2
    int* s_order = reinterpret_cast<int*>(s_mis + kPartitionings);                    // [1024]
    float* s_eu = reinterpret_cast<float*>(s_order + kPartitionings);                 // [1024]
    float* s_es = s_eu + kPartitionings;                                               // [1024]
    int* s_start = reinterpret_cast<int*>(s_es + kPartitionings);                     // [64]
    int* s_hist = s_start + 64;                                                        // [64]
    const int n = *list_count, nsel = T.count_selected[pc - 1], nk = min(tc, 64);
    if (n > T.exw_max_list) return;  // k_partition_ex takes this list
    auto PX = [&](int t, float* o) __attribute__((always_inline)) {
        const float4 f = s_px[t];
        o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = f.w;
    };
    for (int li = blockIdx.x; li < n; li += gridDim.x) {
        const int k = list[li];
        const BlockLoc loc = block_loc(A, T, A.blk0 + k);
        uint32_t amin = 0xFFu, amax = 0u;
        for (int t = lane; t < tc; t += 64) {
            const uint32_t v = texel_rgba8(loc, T, t);
            s_px[t] = make_float4(u8_search_value(v & 0xFF), u8_search_value((v >> 8) & 0xFF),
                                  u8_search_value((v >> 16) & 0xFF), u8_search_value(v >> 24));
            amin = min(amin, v >> 24);
            amax = max(amax, v >> 24);
        }
        for (int m = 32; m > 0; m >>= 1) {
            amin = min(amin, (uint32_t)__shfl_xor((int)amin, m));
            amax = max(amax, (uint32_t)__shfl_xor((int)amax, m));
        }
        if (lane < 64) s_hist[lane] = 0;
        __syncthreads();
        // --- k-means (best_partitionings.cpp:26-194, 419-430): seeding and
        // centre sums in texel order on every lane, assignment per texel
        float centers[4][4];
        {
            float px[4];
            PX(145897 % tc, px);
            for (int c = 0; c < 4; ++c) centers[0][c] = px[c];
        }
#pragma unroll
        for (int q = 1; q < pc; ++q) {
            auto dmin = [&](int t) __attribute__((always_inline)) {
                float px[4];
                PX(t, px);
                float d = ex_dist(px, centers[0]);
#pragma unroll
                for (int j = 1; j < q; ++j) {
                    const float dj = ex_dist(px, centers[j]);
                    d = dj < d ? dj : d;
                }
                return d;
            };
            float dsum = 0.0f;
            for (int t = 0; t < tc; ++t) dsum += dmin(t);
            const float cut = dsum * kKmeansCutoffs[3 * (pc - 2) + q - 1];
            float summa = 0.0f;
            int next = 0;
            for (; next < tc; ++next) {
                summa += dmin(next);
                if (summa >= cut) break;
            }
            next = min(next, tc - 1);
            float px[4];
            PX(next, px);
            for (int c = 0; c < 4; ++c) centers[q][c] = px[c];
        }
#pragma unroll
        for (int q = pc; q < 4; ++q)
            for (int c = 0; c < 4; ++c) centers[q][c] = 0.0f;
        for (int it = 0; it < 3; ++it) {
            if (it > 0) {
                float sum[4][4] = {};
                int cnt[4] = {0, 0, 0, 0};
                for (int t = 0; t < tc; ++t) {
                    float px[4];
                    PX(t, px);
                    const int p = s_part[t];
#pragma unroll
                    for (int q = 0; q < pc; ++q)
                        if (p == q) {
                            for (int c = 0; c < 4; ++c) sum[q][c] += px[c];
                            cnt[q]++;
                        }
                }
#pragma unroll
                for (int q = 0; q < pc; ++q) {
                    const float sc = cnt[q] ? 1.0f / (float)cnt[q] : 0.0f;
                    for (int c = 0; c < 4; ++c) centers[q][c] = cnt[q] ? sum[q][c] * sc : 0.0f;
                }
                __syncthreads();  // every lane has read s_part
            }
            int cnt[4] = {0, 0, 0, 0};
            for (int t0 = 0; t0 < tc; t0 += 64) {
                const int t = t0 + lane;
                int bp = -1;
                if (t < tc) {
                    float px[4];
                    PX(t, px);
                    float bd = 3.402823466e38f;
                    bp = 0;
#pragma unroll
                    for (int q = 0; q < pc; ++q) {
                        const float d = ex_dist(px, centers[q]);
                        if (d < bd) { bd = d; bp = q; }
                    }
                    s_part[t] = (uint8_t)bp;
                }
#pragma unroll
                for (int q = 0; q < pc; ++q) cnt[q] += __popcll(__ballot(bp == q));
            }
            __syncthreads();
            bool problem;
            do {
                problem = false;
#pragma unroll
                for (int q = 0; q < pc; ++q)
                    if (cnt[q] == 0) {
                        problem = true;
                        const int old = s_part[q];
#pragma unroll
                        for (int r = 0; r < pc; ++r) cnt[r] -= r == old;
                        cnt[q]++;
                        __syncthreads();  // every lane has read s_part[q]
                        if (lane == 0) s_part[q] = (uint8_t)q;
                        __syncthreads();
                    }
            } while (problem);
        }
        // --- mismatch of every selected partitioning, counting sort
        // (best_partitionings.cpp:325-397)
        uint64_t a[4] = {0, 0, 0, 0};
        {
            const int p = lane < nk ? s_part[T.kmeans_texels[lane]] : -1;
#pragma unroll
            for (int q = 0; q < pc; ++q) a[q] = __ballot(p == q);
        }
        for (int i = lane; i < nsel; i += 64) {
            const uint64_t* bb = T.coverage + ((size_t)(pc - 2) * kPartitionings + i) * 4;
            auto Pp = [&](int x, int y) { return __popcll(a[x] ^ bb[y]); };
            int v;
            if (pc == 2) {
                v = min(Pp(0, 0) + Pp(1, 1), Pp(0, 1) + Pp(1, 0));
            } else if (pc == 3) {
                const int v0 = min(Pp(1, 1) + Pp(2, 2), Pp(1, 2) + Pp(2, 1)) + Pp(0, 0);
                const int v1 = min(Pp(1, 0) + Pp(2, 2), Pp(1, 2) + Pp(2, 0)) + Pp(0, 1);
                const int v2 = min(Pp(1, 0) + Pp(2, 1), Pp(1, 1) + Pp(2, 0)) + Pp(0, 2);
                v = min(v0, min(v1, v2));
            } else {
                const int mx23 = min(Pp(2, 2) + Pp(3, 3), Pp(2, 3) + Pp(3, 2));
                const int mx13 = min(Pp(2, 1) + Pp(3, 3), Pp(2, 3) + Pp(3, 1));
                const int mx12 = min(Pp(2, 1) + Pp(3, 2), Pp(2, 2) + Pp(3, 1));
                const int mx03 = min(Pp(2, 0) + Pp(3, 3), Pp(2, 3) + Pp(3, 0));
                const int mx02 = min(Pp(2, 0) + Pp(3, 2), Pp(2, 2) + Pp(3, 0));
                const int mx01 = min(Pp(2, 1) + Pp(3, 0), Pp(2, 0) + Pp(3, 1));
                const int v0 = Pp(0, 0) + min(Pp(1, 1) + mx23, min(Pp(1, 2) + mx13, Pp(1, 3) + mx12));
                const int v1 = Pp(0, 1) + min(Pp(1, 0) + mx23, min(Pp(1, 2) + mx03, Pp(1, 3) + mx02));
                const int v2 = Pp(0, 2) + min(Pp(1, 1) + mx03, min(Pp(1, 0) + mx13, Pp(1, 3) + mx01));
                const int v3 = Pp(0, 3) + min(Pp(1, 1) + mx02, min(Pp(1, 2) + mx01, Pp(1, 0) + mx12));
                v = min(v0, min(v1, min(v2, v3)));
            }
            const int m = v / 2;  // < 64
            s_mis[i] = (uint8_t)m;
            atomicAdd(&s_hist[m], 1);
            s_order[i] = 0;
        }
        __syncthreads();
        if (lane == 0) {
            // the running sum stops at the k-means texel count as in the
            // reference; buckets above keep their count as the start
            int sum = 0;
            for (int m = 0; m < 64; ++m) {
                const int c = s_hist[m];
                s_start[m] = m < nk ? sum : c;
                if (m < nk) sum += c;
            }
        }
        __syncthreads();
        // stable placement, 64 partitionings at a time in index order: a
        // lane's slot is its bucket's start plus the lanes of the same
        // bucket before it; a slot written twice keeps the later index
        for (int i0 = 0; i0 < nsel; i0 += 64) {
            const int i = i0 + lane;
            const bool valid = i < nsel;
            const int m = valid ? s_mis[i] : 0;
            uint64_t same = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                const uint64_t bm = __ballot(valid && ((m >> b) & 1));
                same &= ((m >> b) & 1) ? bm : ~bm;
            }
            const int rank = __popcll(same & ((1ull << lane) - 1ull));
            const int pos = valid ? s_start[m] + rank : 0;
            __syncthreads();
            if (valid && rank == __popcll(same) - 1) s_start[m] += __popcll(same);
            if (valid && pos < nsel) atomicMax(&s_order[pos], i);
            __syncthreads();
        }
        // --- line fits of the first search_limit partitionings in that order
        const bool rgba = amin != amax;
        float wie = 0.055f;
        if (tc <= 20) wie = 0.03f;
        else if (tc <= 31) wie = 0.04f;
        else if (tc <= 41) wie = 0.05f;
        wie = wie * wie;
        const int nsearch = min(nsel, search_limit);
        for (int sidx = lane; sidx < nsearch; sidx += 64) {
            float ue, se;
            ex_line_fit<PC, kW>(PX, T, pc, s_order[sidx], rgba, wie, &ue, &se);
            s_eu[sidx] = ue;
            s_es[sidx] = se;
        }
        __syncthreads();
        // --- insert_result in mismatch order (every lane alike)
        float ube[4] = {1e30f, 1e30f, 1e30f, 1e30f}, sbe[4] = {1e30f, 1e30f, 1e30f, 1e30f};
        int ubi[4] = {0, 0, 0, 0}, sbi[4] = {0, 0, 0, 0};
        for (int sidx = 0; sidx < nsearch; ++sidx) {
            const int i = s_order[sidx];
            ex_insert(s_eu[sidx], i, ube, ubi);
            ex_insert(s_es[sidx], i, sbe, sbi);
        }
        const int req = min(nsearch, 4);
        int pos[4] = {-1, -1, -1, -1}, emitted = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int idx = (j & 1) ? sbi[j >> 1] : ubi[j >> 1];
            bool dup = emitted >= req || (j >> 1) >= req;
#pragma unroll
            for (int q = 0; q < 4; ++q) dup |= q < emitted && pos[q] == idx;
            if (!dup) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q == emitted) pos[q] = idx;
                ++emitted;
            }
        }
        if (lane == 0)
            for (int c = 0; c < 4; ++c) pb_part[k * 4 + c] = (int16_t)(pos[c] >= 0 ? pos[c] : pos[0]);
        __syncthreads();  // the next block reuses the LDS
    }
}

// ---- the learned policy's seeded candidates of a short block list ---------
// The seeded search of k_partition_ex (its `seeds` predicted seeds of the
// count, ranked by the uncorrelated / same-chroma line fits), one wave per
// listed block instead of one lane: lane s fits the block's s-th seed (the
// block's texels in LDS, read by every lane alike), then every lane runs the
// top-4 insertion over the seeds in their listed order, so the candidates are
// those of the lane-per-block kernel bit for bit.  A lane per block keeps 48
// line fits in one lane's chain, which bounds the latency of a small image;
// the workgroups loop over the list like k_partition_exw's.  Only without a
// mismatch-order prefix (mismatch_candidates 0); the lane kernel does that.
size_t partition_exs_lds_bytes(const DevTables& T) { return (size_t)T.tc * 16 + 3 * 64 * 4; }

template <int PC, int kW>
__global__ void __launch_bounds__(64) k_partition_exs(ImageArgs A, DevTables T, int16_t* pb_part, const int* list,
                                                      const int* list_count, SearchPlan P) {
    constexpr int pc = PC;
    extern __shared__ __align__(16) uint8_t smem[];
    const int tc = T.tc, lane = threadIdx.x;
    float4* s_px = reinterpret_cast<float4*>(smem);               // [tc]
    float* s_eu = reinterpret_cast<float*>(s_px + tc);           // [64] per seed slot
    float* s_es = s_eu + 64;                                      // [64]
    int* s_i = reinterpret_cast<int*>(s_es + 64);                 // [64] packed entry, -1 = not fitted
    const int n = *list_count, nsel = T.count_selected[pc - 1];
    if (n > T.exw_max_list || P.mis_candidates > 0) return;  // k_partition_ex takes this list
    const int nseeds = min(min(P.seeds, P.kmax), 64);
    auto PX = [&](int t, float* o) __attribute__((always_inline)) {
        const float4 f = s_px[t];
        o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = f.w;
    };
    for (int li = blockIdx.x; li < n; li += gridDim.x) {
        const int k = list[li];
        const BlockLoc loc = block_loc(A, T, A.blk0 + k);
        uint32_t amin = 0xFFu, amax = 0u;
        for (int t = lane; t < tc; t += 64) {
            const uint32_t v = texel_rgba8(loc, T, t);
            s_px[t] = make_float4(u8_search_value(v & 0xFF), u8_search_value((v >> 8) & 0xFF),
                                  u8_search_value((v >> 16) & 0xFF), u8_search_value(v >> 24));
            amin = min(amin, v >> 24);
            amax = max(amax, v >> 24);
        }
        for (int m = 32; m > 0; m >>= 1) {
            amin = min(amin, (uint32_t)__shfl_xor((int)amin, m));
            amax = max(amax, (uint32_t)__shfl_xor((int)amax, m));
        }
        __syncthreads();
        const bool rgba = amin != amax;
        float wie = 0.055f;
        if (tc <= 20) wie = 0.03f;
        else if (tc <= 31) wie = 0.04f;
        else if (tc <= 41) wie = 0.05f;
        wie = wie * wie;
        // the list ends at its first seed outside the head; seeds that are
        // not selected partitionings of this count are skipped
        const int* seeds = P.topk_idx + (size_t)k * 5 * P.kmax + (size_t)(pc - 1) * P.kmax;
        const int seed = lane < nseeds ? seeds[lane] : -1;
        const uint64_t ends = __ballot(lane < nseeds && (seed < 0 || seed >= kPartitionings));
        const int nlist = ends ? (__ffsll((unsigned long long)ends) - 1) : nseeds;
        int i = -1;
        if (lane < nlist) {
            i = T.seed_to_packed[(pc - 2) * kPartitionings + seed];
            if (i >= nsel) i = -1;
        }
        if (i >= 0) {
            float ue, se;
            ex_line_fit<PC, kW>(PX, T, pc, i, rgba, wie, &ue, &se);
            s_eu[lane] = ue;
            s_es[lane] = se;
        }
        s_i[lane] = i;
        __syncthreads();
        float ube[4] = {1e30f, 1e30f, 1e30f, 1e30f}, sbe[4] = {1e30f, 1e30f, 1e30f, 1e30f};
        int ubi[4] = {0, 0, 0, 0}, sbi[4] = {0, 0, 0, 0};
        int nvalid = 0;
        for (int q = 0; q < nlist; ++q) {
            const int iq = s_i[q];
            if (iq < 0) continue;
            ++nvalid;
            ex_insert(s_eu[q], iq, ube, ubi);
            ex_insert(s_es[q], iq, sbe, sbi);
        }
        const int req = min(nvalid, 4);
        int pos[4] = {-1, -1, -1, -1}, emitted = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int idx = (j & 1) ? sbi[j >> 1] : ubi[j >> 1];
            bool dup = emitted >= req || (j >> 1) >= req;
#pragma unroll
            for (int q = 0; q < 4; ++q) dup |= q < emitted && pos[q] == idx;
            if (!dup) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q == emitted) pos[q] = idx;
                ++emitted;
            }
        }
        if (lane < 4) pb_part[k * 4 + lane] = (int16_t)(lane == 0 ? pos[0] : lane == 1 ? pos[1] : lane == 2 ? pos[2] : pos[3]);
        __syncthreads();  // the next block reuses the LDS
    }
}

void launch_partition_ex(bool seeded, int pc, const DevTables& T, dim3 grid, size_t lds, hipStream_t s,
                         const ImageArgs& A, int16_t* pb_part, uint8_t* scratch, const int* list, const int* list_count,
                         int search_limit, const SearchPlan& P) {
#define ASTCG_PEX(PCV, F4, SD)                                                                                       \
    hipLaunchKernelGGL((k_partition_ex<PCV, F4, SD>), grid, dim3(64), lds, s, A, T, pb_part, scratch, list, list_count, \
                       search_limit, P)
#define ASTCG_PEX_PC(SD)                                                 \
    if (pc == 2) { if (f4) ASTCG_PEX(2, true, SD); else ASTCG_PEX(2, false, SD); } \
    else if (pc == 3) { if (f4) ASTCG_PEX(3, true, SD); else ASTCG_PEX(3, false, SD); } \
    else { if (f4) ASTCG_PEX(4, true, SD); else ASTCG_PEX(4, false, SD); }
    const bool f4 = T.tc <= T.ex_f4_max_tc;
    if (seeded) { ASTCG_PEX_PC(true) }
    else { ASTCG_PEX_PC(false) }
#undef ASTCG_PEX_PC
#undef ASTCG_PEX
}

hipError_t launch_chunk(const DevTables& T, const ImageArgs& A_in, const ChunkBuffers& B, SymBest* best, int max_pc,
                        hipStream_t s, Profiler* prof, const LabelBufs* labels, const SearchPlan* plan,
                        bool dual_plane, int exhaustive) {
    (void)hipGetLastError();  // the final check reports this call's launches only
    ImageArgs A0 = A_in;
    A0.blk_texels = B.texels;
    {
        const size_t n = (size_t)A0.nb * T.tc;
        hipLaunchKernelGGL(k_block_texels, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A0, T, B.texels);
    }
    const LabelBufs lab = labels ? *labels : LabelBufs{nullptr, nullptr, nullptr};
    const int threads = 64;
    auto blocks_for = [&](long long n) { return (unsigned)((n + threads - 1) / threads); };
    const uint32_t* mode_mask = nullptr;
    const uint8_t* slots = nullptr;
    int mask_words = 0;
    if (plan) {
        hipLaunchKernelGGL(k_plan_zero, dim3(1), dim3(64), 0, s, plan->pb_count);
        hipLaunchKernelGGL(k_plan_modes, dim3(blocks_for(A0.nb)), dim3(64), 0, s, A0, T, *plan, T.nbm,
                           const_cast<uint32_t*>(plan->mode_mask), const_cast<uint8_t*>(plan->slots));
        mode_mask = plan->mode_mask;
        slots = plan->slots;
        mask_words = plan->mask_words;
    }
    for (int pc = 1; pc <= max_pc; ++pc) {
        const int npos = pc == 1 ? 1 : 4;
        const int npb = A0.nb * npos;
        ImageArgs A = A0;
        if (pc > 1) {
            if (prof) prof->begin(K_PARTITION, s);
            if (exhaustive) {
                // the mode-error buffer of this count is free until k_dm_bm;
                // under the learned policy only the blocks that search this
                // count run, compacted into a list (the plan's partitioned-block
                // list, filled by k_plan only afterwards)
                const dim3 pgrid((A.nb + 63) / 64);
                const size_t plds = partition_ex_lds_bytes(T);
                uint8_t* scr = reinterpret_cast<uint8_t*>(B.ccr);
                int* list = nullptr;
                int* list_count = nullptr;
                if (plan) {
                    list = plan->pb_list;
                    list_count = plan->pb_count + 3;
                    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(256), 0, s, list_count, 1);
                    hipLaunchKernelGGL(k_need_list, dim3((A.nb + kNeedListThreads - 1) / kNeedListThreads), dim3(kNeedListThreads), 0, s, A,
                                       *plan, pc, 0, list, list_count);
                }
                if (plan) {
                    // hard and fallback blocks: a wave per block for a short
                    // list, a lane per block for a long one (each kernel
                    // reads the list length and only one of them works)
                    launch_partition_ex(false, pc, T, pgrid, plds, s, A, B.pb_part, scr, list, list_count, exhaustive,
                                        SearchPlan{});
                    const dim3 wgrid((unsigned)std::min(A.nb, kExwGrid));
                    const size_t wlds = partition_exw_lds_bytes(T);
#define ASTCG_PEXW(PCV)                                                                                              \
    if (T.tc <= 64) hipLaunchKernelGGL((k_partition_exw<PCV, 1>), wgrid, dim3(64), wlds, s, A, T, B.pb_part, list,  \
                                       list_count, exhaustive);                                                     \
    else hipLaunchKernelGGL((k_partition_exw<PCV, (kMaxTexels + 63) / 64>), wgrid, dim3(64), wlds, s, A, T,         \
                            B.pb_part, list, list_count, exhaustive)
                    if (pc == 2) { ASTCG_PEXW(2); }
                    else if (pc == 3) { ASTCG_PEXW(3); }
                    else { ASTCG_PEXW(4); }
#undef ASTCG_PEXW
                } else {
                    launch_partition_ex(false, pc, T, pgrid, plds, s, A, B.pb_part, scr, list, list_count, exhaustive,
                                        SearchPlan{});
                }
            } else {
                const dim3 pgrid((A.nb + kPartBlocks - 1) / kPartBlocks);
                const size_t plds = partition_lds_bytes(T);
                const SearchPlan pplan = plan ? *plan : SearchPlan{};
                const int use = plan ? 1 : 0;
                if (pc == 2) hipLaunchKernelGGL(k_partition<2>, pgrid, dim3(64), plds, s, A, T, B.pb_part, pplan, use);
                else if (pc == 3) hipLaunchKernelGGL(k_partition<3>, pgrid, dim3(64), plds, s, A, T, B.pb_part, pplan, use);
                else hipLaunchKernelGGL(k_partition<4>, pgrid, dim3(64), plds, s, A, T, B.pb_part, pplan, use);
            }
            if (plan && plan->seeds > 0 && (plan->seed_topk || plan->seeds > 4)) {
                // blocks searching this count from their predicted seeds (a
                // second list after the first): the seed head's top-k of
                // those blocks, then (more than 4 seeds) their line fits rank them
                const dim3 pgrid((A.nb + 63) / 64);
                int* list = plan->pb_list + A.nb;
                int* list_count = plan->pb_count + 4;
                hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(256), 0, s, list_count, 1);
                hipLaunchKernelGGL(k_need_list, dim3((A.nb + kNeedListThreads - 1) / kNeedListThreads), dim3(kNeedListThreads), 0, s, A,
                                   *plan, pc, 1, list, list_count);
                if (plan->seed_topk &&
                    astcg_mlp_seed_topk(list, list_count, A.nb, plan->h2, plan->w3_seed + (size_t)(pc - 2) * 1024 * plan->hidden,
                                        plan->b3_seed + (pc - 2) * 1024, plan->seeds, const_cast<int*>(plan->topk_idx),
                                        5 * plan->kmax, (pc - 1) * plan->kmax, s) != ASTCG_OK)
                    return hipErrorLaunchFailure;
                if (plan->seeds > 4) {
                    // a lane per block for a long list, a wave per block for a
                    // short one (each kernel reads the list length)
                    launch_partition_ex(true, pc, T, pgrid, partition_ex_lds_bytes(T), s, A, B.pb_part,
                                        reinterpret_cast<uint8_t*>(B.ccr), list, list_count, 0, *plan);
                    const dim3 sgrid((unsigned)std::min(A.nb, kExwGrid));
                    const size_t slds = partition_exs_lds_bytes(T);
#define ASTCG_PEXS(PCV)                                                                                              \
    if (T.tc <= 64) hipLaunchKernelGGL((k_partition_exs<PCV, 1>), sgrid, dim3(64), slds, s, A, T, B.pb_part, list,  \
                                       list_count, *plan);                                                          \
    else hipLaunchKernelGGL((k_partition_exs<PCV, (kMaxTexels + 63) / 64>), sgrid, dim3(64), slds, s, A, T,         \
                            B.pb_part, list, list_count, *plan)
                    if (pc == 2) { ASTCG_PEXS(2); }
                    else if (pc == 3) { ASTCG_PEXS(3); }
                    else { ASTCG_PEXS(4); }
#undef ASTCG_PEXS
                }
            }
            if (plan) {
                hipLaunchKernelGGL(k_plan, dim3(blocks_for(A.nb)), dim3(64), 0, s, A, T, pc, B.pb_part, *plan,
                                   plan->pb_count + (pc - 2));
                A.pb_list = plan->pb_list;
                A.pb_inv = plan->pb_inv;
                A.pb_count = plan->pb_count + (pc - 2);
            }
            if (prof) prof->end(K_PARTITION, s);
        }
        {
            if (prof) prof->begin(K_IDEAL_COLOR, s);
            const dim3 igrid(blocks_for(npb));
#define ASTCG_IDEAL(PCV)                                                                                          \
    hipLaunchKernelGGL(k_ideal_color<PCV>, igrid, dim3(threads), 0, s, A, T, B.pb_part, npb, B.ideal, B.wts, B.ct_err, \
                       B.ct_fmt)
            switch (pc) {
                case 1: ASTCG_IDEAL(1); break;
                case 2: ASTCG_IDEAL(2); break;
                case 3: ASTCG_IDEAL(3); break;
                default: ASTCG_IDEAL(4); break;
            }
#undef ASTCG_IDEAL
            if (prof) prof->end(K_IDEAL_COLOR, s);
        }
        {
            if (prof) prof->begin(K_DM_BM, s);
            for (int g = 0; g < kDmGroups; ++g) {
                const int first = T.dm_group_start[g], count = T.dm_group_start[g + 1] - first;
                if (count <= 0) continue;
                const int w = T.dm_group_w[g];
#define ASTCG_DM_BM(PCV)                                                                                          \
    hipLaunchKernelGGL(k_dm_bm<PCV>, dim3(dm_bm_grid(npb, count)), dim3(64), dm_bm_lds_bytes(T, w, pc), s, A, T, \
                       B.pb_part, npb, B.ideal, B.wts, B.ct_err, B.ct_fmt, B.ccr, B.qw, w, mode_mask, mask_words, first,  \
                       count)
                switch (pc) {
                    case 1: ASTCG_DM_BM(1); break;
                    case 2: ASTCG_DM_BM(2); break;
                    case 3: ASTCG_DM_BM(3); break;
                    default: ASTCG_DM_BM(4); break;
                }
#undef ASTCG_DM_BM
            }
            if (prof) prof->end(K_DM_BM, s);
        }
        {
            if (prof) prof->begin(K_TOP, s);
            float* pb_est = (plan && pc > 1 && plan->keep_refine < 4) ? plan->pb_est : nullptr;
            if (pb_est)
                hipLaunchKernelGGL(k_pb_est, dim3(blocks_for(npb)), dim3(64), 0, s, A, pc, npb, T.nbm, B.ccr,
                                   mode_mask, mask_words, pb_est);
            const int nbins = T.nbm * kRefineClasses;
            const size_t hist = (size_t)nbins * sizeof(int);
            const unsigned tgrid = (unsigned)((npb + kTopThreads - 1) / kTopThreads);
            hipLaunchKernelGGL(k_top, dim3(tgrid), dim3(kTopThreads), hist, s, A, pc, npb, T.nbm, B.ccr, B.top,
                               B.bin_count, B.cand, B.cand_err, mode_mask, mask_words, slots, pb_est,
                               plan ? plan->keep_refine : 4, plan ? plan->fallback : nullptr,
                               T.refine_classes);
            hipLaunchKernelGGL(k_bins, dim3(1), dim3(64), 0, s, nbins, B.bin_count, B.bin_cur, B.work, B.work_total,
                               B.tile_next);
            const int sthreads = T.tc <= T.refine_stage_max_tc ? kTopThreads : 64;
            hipLaunchKernelGGL(k_scatter, dim3((unsigned)((npb + sthreads - 1) / sthreads)), dim3(sthreads), 2 * hist,
                               s, A, pc, npb, nbins, B.top,
                               B.bin_cur, B.work);
            if (prof) prof->end(K_TOP, s);
        }
        {
            if (prof) prof->begin(K_REFINE, s);
            const size_t lds = refine_lds_bytes(T, B.wstride, pc);
            const unsigned tiles = blocks_for((long long)npb * 8 + (long long)T.nbm * kRefineClasses * 63);
            launch_refine(T.tc <= T.refine_stage_max_tc, pc, tiles, lds, s, A, T, B, npb,
                          plan ? plan->iterations : 6);
            if (prof) prof->end(K_REFINE, s);
        }
        ASTCG_LAUNCH(K_PICK, k_pick, blocks_for(A.nb), A, T, pc, B.pb_part, B.cand, B.cand_err, best, lab);
    }
    if (dual_plane && T.ndual > 0 && T.tc <= kDualMaxTc) {
        if (prof) prof->begin(K_PICK, s);
        hipLaunchKernelGGL(k_dual, dim3(blocks_for(A0.nb)), dim3(64), dual_lds_bytes(T), s, A0, T, best);
        if (prof) prof->end(K_PICK, s);
    }
    return hipGetLastError();
}

hipError_t launch_decode(const DevTables& T, const uint8_t* blocks, int width, int height, uint8_t* rgba,
                         hipStream_t s) {
    (void)hipGetLastError();  // the final check reports this call's launches only
    const int blocks_x = (width + T.bx - 1) / T.bx;
    const int nblocks = blocks_x * ((height + T.by - 1) / T.by);
    hipLaunchKernelGGL(k_decode, dim3((nblocks + 63) / 64), dim3(64), 0, s, T, blocks, width, height, blocks_x, nblocks,
                       rgba);
    return hipGetLastError();
}

hipError_t launch_zero_ints(int* p, int n, hipStream_t s) {
    (void)hipGetLastError();  // the final check reports this call's launches only
    if (n > 0) hipLaunchKernelGGL(k_zero_ints, dim3((unsigned)((n + 255) / 256 < 64 ? (n + 255) / 256 : 64)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_fb_reset(unsigned long long* fb_count, hipStream_t s) {
    (void)hipGetLastError();  // the final check reports this call's launches only
    hipLaunchKernelGGL(k_fb_reset, dim3(1), dim3(64), 0, s, fb_count);
    return hipGetLastError();
}

hipError_t launch_pack(const DevTables& T, const ImageArgs& A, const SymBest* best, hipStream_t s, Profiler* prof,
                       bool void_extent) {
    (void)hipGetLastError();  // the final check reports this call's launches only
    const int threads = 64;
    ASTCG_LAUNCH(K_PACK, k_pack, (A.num_blocks + threads - 1) / threads, T, A, best, void_extent ? 1 : 0);
    return hipGetLastError();
}

}  // namespace astcg
