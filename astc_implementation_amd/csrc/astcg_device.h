// Device-side data layout shared by the host orchestration and the gfx950
// kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include <cstddef>

#include "astcg_tables.h"

namespace astcg {

// Read-only per-footprint tables in HBM (one allocation, see upload_tables).
constexpr int kDmGroups = 4;  // weight-count groups of k_dm_bm's launches

struct DevTables {
    const DmInfo* dm;
    const BmInfo* bm;
    const uint16_t* dm_bm_off;    // [ndm + 1] CSR over block modes grouped by decimation mode
    const uint16_t* dm_bm_list;   // [nbm] block-mode trial indices
    const uint8_t* t2w_count;     // [ndm][tc]
    const uint8_t* t2w_idx;       // [ndm][tc][4]
    const float* t2w_w;           // [ndm][tc][4]
    const uint8_t* w2t_count;     // [ndm][64]
    const uint16_t* w2t_off;      // [ndm][64]
    const uint8_t* w2t_idx;       // pool
    const float* w2t_w;           // pool
    const uint8_t* part_of_texel; // [3][1024][tc]
    const uint16_t* part_seed;    // [3][1024]
    const uint8_t* part_counts;   // [3][1024][4]
    const uint64_t* coverage;     // [3][1024][4]
    const uint16_t* ordering;     // [3][2][128]
    const uint32_t* kmeans_texels;// [64]
    const float* sin_t;           // [1024][16]
    const float* cos_t;           // [1024][16]
    const int16_t* raw_to_bm;     // [2048]
    const uint8_t* raw_wc;        // [2048]
    const uint8_t* raw_quant;     // [2048]
    const uint16_t* seed_to_packed; // [3][1024] partition seed -> packed table entry
    const uint16_t* qw_off;       // [nbm] start of a trial mode's weight region, in units of npb bytes
    int qw_stride;                // bytes of packed quantised weights per partitioned block
    const DualMode* dual;         // [ndual] dual-plane modes of the optional dual-plane search
    int ndual;
    int bx, by, tc, ndm, nbm;
    int count_selected[4];
    // k_dm_bm launches one group of decimation modes at a time, grouped by
    // weight count (<= 16, <= 24, <= 32, more), each with LDS rows sized for
    // its widest grid: dm_order lists the modes group by group
    const uint8_t* dm_order;      // [ndm]
    int dm_group_start[kDmGroups + 1];  // group g = dm_order[dm_group_start[g] .. dm_group_start[g + 1])
    int dm_group_w[kDmGroups];          // LDS row (weights, multiple of 4) of group g
    // launch shapes (astcg_tuning), none of which changes a result
    int refine_stage_max_tc;      // k_refine stages the block's RGBA8 texels in LDS up to this texel count
    int ex_f4_max_tc;             // k_partition_ex stages float texels in LDS up to this count (<= 64)
    int exw_max_list;             // block lists up to this length run k_partition_exw / k_partition_exs
    int refine_queue;             // 0: k_refine, a workgroup per work-list tile; n >= 1: the persistent
                                  // k_refine_q, whose waves refill once n of their lanes are free
    int refine_classes;           // 1: the work list is bucketed by trial mode and format class; 0: by mode
};

// Launch-shape defaults (measured on the 4096x4096 benchmark images).
constexpr int kRefineStageMaxTc = 25;  // larger blocks ran faster unstaged, with more workgroups per CU
constexpr int kExF4MaxTc = 16;         // keeps 6 k_partition_ex workgroups per CU
constexpr int kExwMaxList = 65536;     // lane-per-block throughput is higher once its waves fill the chip
constexpr int kExwAllListsTc = 36;     // from this many texels per block every list runs a wave per block
constexpr int kExwAllLists = 1 << 20;
constexpr int kRefineQueue = 0;        // the persistent refinement measured slower on the 4096^2 images

// One texture of a launch: its blocks occupy [first_block, first_block +
// num_blocks) of the launch's block index space; `out` receives its 16-byte
// blocks in row-major block order.
struct TexDesc {
    const uint8_t* rgba;
    uint8_t* out;
    int width, height, blocks_x;
    int first_block, num_blocks;
    int pad;
};

// Image and chunk geometry.  A single image is one texture held by value;
// a batch keeps its texture table (sorted by first_block) in HBM.
struct ImageArgs {
    TexDesc t0;               // the texture when ntex == 1
    const TexDesc* tex;       // [ntex] when ntex > 1
    int ntex;
    int num_blocks;           // blocks of all textures
    int blk0;                 // first block of the chunk
    int nb;                   // blocks in the chunk
    // partitioned blocks of the current partition count (learned policy;
    // null = all nb * 4 positions): entry = block * 4 + position
    const int* pb_list;
    const int* pb_inv;        // [nb * 4] list index of each position, -1 = not searched
    const int* pb_count;      // entries in pb_list
    // [nb][texel count] RGBA8 texels of the chunk's blocks, edge-clamped,
    // one contiguous row per block (k_block_texels; set by launch_chunk)
    const uint32_t* blk_texels;
};

// Search outcome recorded for training the candidate predictor (all over
// the blocks of one launch chain, num_blocks of ImageArgs): per partition
// count 2..4 the seed of the best partitioned block, and the seed and best
// error of each of the 4 compat positions (-1 / 1e30 where not searched).
struct LabelBufs {
    int16_t* best_seed;  // [3][num_blocks]
    int16_t* cand_seed;  // [3][4][num_blocks]
    float* cand_err;     // [3][4][num_blocks]
};

// Candidate plan of the learned policy for one chunk (null: the plain
// search, every block's 4 partition positions, all modes, 8 slots).
struct SymBest;

struct SearchPlan {
    const uint32_t* mode_mask;  // [nb][mask_words] trial modes each block may use
    int mask_words;
    const uint8_t* slots;       // [nb] candidates refined per partitioned block (1..8)
    // builds the partition-count lists after k_partition (see launch_plan)
    int* pb_list;               // [nb * 4]
    int* pb_inv;                // [nb * 4]
    int* pb_count;              // [3]
    const int* topk_idx;
    const float* topk_prob;
    int kmax, seeds, modes, nslots;
    float fallback_below, count_below;
    uint8_t* fallback;          // [nb]
    unsigned long long* fb_count;  // running count of fallback blocks (one counter per encode call)
    // ranking of the compat positions (keep_pb < 4): the blocks' second
    // hidden activations and the seed-head weights, both bf16
    int keep_pb;                // positions kept per searched count (1..4)
    int hidden;
    const uint16_t* h2;         // [nb][hidden]
    const uint16_t* w3_seed;    // [3][1024][hidden] (seed-head columns, transposed)
    const float* b3_seed;       // [3][1024]
    // the seed heads' top-k is computed per partition count for the blocks
    // that search it from their seeds (from h2), not with the other heads
    int seed_topk;
    // refinement pruning (keep_refine < 4): each partitioned block's best
    // estimated candidate error
    int keep_refine;
    float* pb_est;              // [nb * 4]
    // early out: blocks whose best error so far is below error_limit skip
    // the remaining partition counts (0 = off)
    float error_limit;
    // blocks whose best error after one partition is above full_limit get
    // the full search for 2..4 partitions (0 = off)
    float full_limit;
    // blocks whose best error so far is above count_limit search every
    // partition count (0 = off)
    float count_limit;
    int hard_keeps_modes;       // hard blocks keep their predicted modes and slots
    int mis_candidates;         // seeded blocks also line-fit this many of the mismatch order first
    float line_limit;           // blocks whose best-line error is above this fall back (0 = off)
    int iterations;             // refinement rounds per candidate (1..6)
    int part_slots;             // slots of non-fallback blocks for 2..4 partitions (1..8)
    const SymBest* best;        // running best per block (all blocks of the launch chain)
};

// pass01 output (ideal endpoints), per partitioned block; texel weights are
// stored separately ([pb][tc]).
struct Ideal {
    float avg[4][4];
    float ep0[4][4];
    float ep1[4][4];
    float wes[4];        // weight error scale per partition
    float cutoff;        // min_weight_cutoff
    float pad[3];
};

// pass09 (1 partition: [quant][int count]) or pass10 (2-4 partitions:
// [quant][combined int count]) colour-format table: error and 4 x 4-bit
// formats per (quant 4..20, column 0..12, partitioned block), stored
// [quant - 4][column][PB].
constexpr int kColorTabRows = 17 * 13;

// pass11 result per (partitioned block, trial block mode).
struct Ccr {
    float total_error;
    int8_t ql, ql_mod;
    uint16_t formats;  // 4 x 4 bits
};

// The refinement's work list is bucketed by (trial block mode, class): the
// class groups candidates whose endpoint packing (pass14) takes the same
// format path, so the lanes of a wave branch alike.  Bucket key = mode *
// kRefineClasses + class.
constexpr int kRefineClassBits = 4;
constexpr int kRefineClasses = 1 << kRefineClassBits;

// Best refined candidate per (partitioned block, candidate slot).
struct CandOut {
    float error;
    uint16_t bm;       // trial block mode
    uint8_t matched;
    uint8_t quant;     // final colour quant level
    uint16_t formats;  // 4 x 4 bits
    uint16_t pad;
    uint8_t packed[32];  // packed and qw are copied as one 96-byte run of dwords
    uint8_t qw[64];
};
static_assert(offsetof(CandOut, packed) % 4 == 0 && offsetof(CandOut, qw) == offsetof(CandOut, packed) + 32,
              "CandOut packed/qw layout");

// Best symbolic block so far, per image block.
struct SymBest {
    float error;
    uint16_t raw_mode;
    uint16_t seed;
    uint8_t pc, matched, quant;
    uint8_t pad;  // dual plane (optional search): plane-2 channel + 1, 0 = single plane
    uint16_t formats;
    uint16_t pad2;
    uint8_t packed[32];
    uint8_t qw[64];
};
static_assert(offsetof(SymBest, packed) % 4 == 0 && offsetof(SymBest, qw) == offsetof(SymBest, packed) + 32,
              "SymBest packed/qw layout");

struct ChunkBuffers {
    int16_t* pb_part;   // [4 * max_nb]
    Ideal* ideal;       // [4 * max_nb]
    float* wts;         // [tc][npb of the launch]
    float* ct_err;      // [17][13][npb of the launch]
    uint16_t* ct_fmt;   // [17][13][npb of the launch]
    Ccr* ccr;           // [nbm][npb of the launch]
    uint8_t* qw;        // per trial mode a region at qw_off[mode] * npb: [npb][weight count rounded up to 4]
    int16_t* top;       // [4 * max_nb][8] bucket key (trial mode * kRefineClasses + class) of each slot, -1 = none
    CandOut* cand;      // [4 * max_nb][8]
    float* cand_err;    // [4 * max_nb][8] each slot's CandOut::error again, packed for k_pick
    int* bin_count;     // [nbm * kRefineClasses] candidates per bucket (zero between launches)
    int* bin_cur;       // [nbm * kRefineClasses] next free work-list position per bucket
    int* work_total;    // [1] work-list length including padding
    int* tile_next;     // [1] next work-list tile of the persistent refinement's queue (zeroed by k_bins)
    int* work;          // [4 * max_nb * 8 + nbm * kRefineClasses * 64] (partitioned block * 8 + slot), bucketed
    uint32_t* texels;   // [max_nb][texel count] the chunk's block texels (ImageArgs::blk_texels)
    int wstride;        // LDS row of quantised weights: max weight count of the trial grids, multiple of 4
};

// Optional per-kernel timing with HIP events recorded around each launch.
enum KernelId { K_PARTITION = 0, K_IDEAL_COLOR, K_DM_BM, K_TOP, K_REFINE, K_PICK, K_PACK, K_COUNT };
const char* kernel_name(int id);

struct Profiler {
    virtual void begin(int kernel, hipStream_t s) = 0;
    virtual void end(int kernel, hipStream_t s) = 0;
    virtual ~Profiler() = default;
};

void upload_quant_tables(hipStream_t s);
// pc_seed (optional, [3][A.num_blocks]): seed of each partition count's best
// partitioning per block, -1 when that count produced no candidate.
// exhaustive > 0: partition candidates from the exhaustive search
// (k_partition_ex) of the first `exhaustive` partitionings in k-means
// mismatch order (a value >= 1024 covers every selected partitioning)
// instead of the reference's live passes 001-007 (k_partition).
hipError_t launch_chunk(const DevTables& T, const ImageArgs& A, const ChunkBuffers& B, SymBest* best, int max_pc,
                        hipStream_t s, Profiler* prof, const LabelBufs* labels = nullptr,
                        const SearchPlan* plan = nullptr, bool dual_plane = false, int exhaustive = 0);
// ASTC LDR decode of a width x height image's blocks (row-major) to RGBA8.
hipError_t launch_decode(const DevTables& T, const uint8_t* blocks, int width, int height, uint8_t* rgba,
                         hipStream_t s);
// Zeroes the fallback counter of the learned policy (a kernel, so that it
// can be part of a captured graph).
hipError_t launch_fb_reset(unsigned long long* fb_count, hipStream_t s);
// Zeroes n ints with a kernel (graph-capturable, unlike a memset node).
hipError_t launch_zero_ints(int* p, int n, hipStream_t s);

hipError_t launch_pack(const DevTables& T, const ImageArgs& A, const SymBest* best, hipStream_t s, Profiler* prof,
                       bool void_extent);

}  // namespace astcg
