/*
 * CPU oracle for the ASTC LDR block encoder.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference encoder (Jure-Z/ASTC_implementation: the C++ sources under code/ and the 30 WGSL
 * passes under shaders/), written to check the HIP product path.  It is
 * linked or loaded only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product never calls it.
 *
 * Execution semantics the WGSL leaves open are fixed here (and identically
 * in the HIP kernels):
 *   - float atomics (CAS add loops) are summed sequentially in texel/weight
 *     index order;
 *   - min()/max()/clamp() follow WGSL: a NaN operand yields the other one
 *     (C fminf/fmaxf);
 *   - round() is round-half-to-even (rintf);
 *   - f32 -> i32/u32 conversions saturate and map NaN to the type maximum
 *     (Tint's conversion polyfill);
 *   - dot() sums left to right, normalize(v) = v / sqrt(dot(v, v)), no FMA
 *     contraction anywhere;
 *   - atan2 is a fixed single-precision method (Cephes atanf reduction) so
 *     that host and device agree bit for bit;
 *   - out-of-bounds buffer accesses in the reference (pass05/06 stride 13 vs
 *     8 floats per trial, pass005/006/007 stride 512 vs 128) read 0 and drop
 *     writes;
 *   - the pass004 output race resolves as if workgroups ran in index order;
 *   - pass12 slots that receive no candidate are not refined (the reference
 *     would refine stale buffer contents); oracle_stats counts such blocks.
 */
#ifndef ASTC_ORACLE_H
#define ASTC_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_TEXELS 144
#define OR_MAX_WEIGHTS 64
#define OR_MAX_DM 87
#define OR_MAX_BM 2048
#define OR_PARTITIONINGS 1024

/* SymbolicBlock, astc.h:639-655 (same field order and meaning). */
typedef struct {
    float errorval;
    uint32_t block_mode_index;   /* raw 11-bit block mode */
    uint32_t partition_count;
    uint32_t partition_index;    /* 10-bit partition seed */
    uint32_t partition_formats_matched;
    uint32_t quant_mode;         /* colour endpoint quant level */
    uint32_t _padding1;
    uint32_t _padding2;
    uint32_t partition_formats[4];
    uint32_t packed_color_values[32]; /* 0..255 unquantised values */
    uint32_t quantized_weights[64];   /* 0..64 unquantised weights */
} or_symbolic;

typedef struct {
    int max_partitions;   /* 1..4, default 4 */
    int batch_size;       /* reference batch geometry, default 920 */
    int threads;          /* OpenMP threads, 0 = 1 */
    int first_block;      /* encode only blocks [first_block, first_block+block_count) */
    int block_count;      /* 0 = all */
    int policy;           /* partition candidates: 0 = the live passes 001-007 (compat),
                             2 = the exhaustive CPU search of best_partitionings.cpp,
                             3 = seed lists (seeded_positions) */
    int search_limit;     /* policy 2: partitionings line-fitted, in mismatch order (0 = all) */
    const int32_t* seed_list; /* policy 3: every block's partition seeds for 2, 3, 4 partitions, */
    int seed_list_len;        /* seed_list_len each, best first (the learned policy's seed heads) */
    int mismatch_candidates;  /* policy 3, more than 4 seeds: mismatch-order partitionings fitted first */
} or_params;

typedef struct {
    uint64_t blocks;
    uint64_t partitioned_blocks;
    uint64_t short_candidate_lists; /* pass12 produced < 8 candidates */
} or_stats;

/* Whole-image encode (ASTCEncoder::encode, compression_pipeline.cpp:323-535). */
int oracle_encode(const uint8_t* rgba, int width, int height, int bx, int by,
                  const or_params* prm, uint8_t* out_blocks, or_symbolic* out_sym,
                  or_stats* stats);

/* symbolic_to_physical, physical_compression.cpp:843-1003. */
int oracle_symbolic_to_physical(int bx, int by, const or_symbolic* sb, uint8_t out[16]);

/* encode_ise, physical_compression.cpp:672-826. */
void oracle_encode_ise(int quant, int count, const uint8_t* in, uint8_t* out, int bit_offset);

/* get_ise_sequence_bitcount, physical_compression.cpp:625-639. */
int oracle_ise_bitcount(int count, int quant);

/* Table dump for golden checks.  id selects the table; returns the number of
 * int32 entries written (or needed when out == NULL). */
int oracle_table(int id, int32_t* out, int cap);

/* Footprint metadata summary: counts[0..] = {decimation modes, block modes,
 * valid dm, valid bm, selected p2, selected p3, selected p4, texel->weight
 * entries} */
int oracle_footprint_info(int bx, int by, int32_t* counts, int cap);

/* Per-footprint metadata dump: partition_of_texel for packed entry idx of
 * partition count pc (pc = 1..4); returns texel count. */
int oracle_partition_entry(int bx, int by, int pc, int idx, uint8_t* part_of_texel,
                           int32_t* seed_and_counts);

/* Decimation table of trial decimation mode m (see or_tables.c). */
int oracle_decimation_entry(int bx, int by, int m, int32_t* out, int cap);

/* The exhaustive partition candidates (code/best_partitionings.cpp:498-737)
 * of block block_index for partition count pc: 4 packed table entries
 * (-1 where fewer were found). */
int oracle_debug_exhaustive_candidates(const uint8_t* rgba, int width, int height, int bx, int by,
                                       int block_index, int partition_count, int32_t* entries_out);
/* The 4 packed table entries (-1 = none) of the seeded candidates
 * (seeded_positions) of one block for the seed list seeds[0 .. nseeds),
 * after mismatch_candidates partitionings of the mismatch order. */
int oracle_debug_seeded_candidates(const uint8_t* rgba, int width, int height, int bx, int by, int block_index,
                                   int partition_count, const int32_t* seeds, int nseeds, int mismatch_candidates,
                                   int32_t* entries_out);

/* Stage-level hooks used by unit tests (single partitioned block). */
int oracle_debug_kmeans_candidates(const uint8_t* rgba, int width, int height, int bx, int by,
                                   int block_index, int partition_count, int batch_size,
                                   int32_t* positions_out);

#ifdef __cplusplus
}
#endif
#endif
