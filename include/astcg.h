/*
 * astcg: ASTC LDR 2D block encoder for AMD Instinct MI355X (gfx950).
 *
 * C ABI that replaces the reference's C++ entry points
 * (Jure-Z/ASTC_implementation):
 *   ASTCEncoder::ASTCEncoder / init      code/astc.h:724-730,  compression_pipeline.cpp:200-233
 *   ASTCEncoder::secondaryInit           code/astc.h:731,      compression_pipeline.cpp:263-321
 *   ASTCEncoder::encode                  code/astc.h:733,      compression_pipeline.cpp:323-535
 *   symbolic_to_physical                 code/astc.h:718-722,  physical_compression.cpp:843-1003
 *   encode_ise                           code/astc.h:710-716,  physical_compression.cpp:672-826
 *   get_ise_sequence_bitcount            code/astc.h:676,      physical_compression.cpp:625-639
 *   store_image                          code/astc_store.h:8-16, astc_store.cpp:18-61
 *   create_astc_file_in_memory           code/astc_store.h:18-30, astc_store.cpp:65-104
 *
 * Conventions: every function returns an astcg_status; nothing throws across
 * the ABI.  The caller owns all host buffers.  One encoder per thread; calls
 * on distinct encoders are independent.
 */
#ifndef ASTCG_H
#define ASTCG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ASTCG_OK = 0,
    ASTCG_ERR_INVALID_ARGUMENT = 1,   /* null pointer, zero size, bad enum */
    ASTCG_ERR_INVALID_FOOTPRINT = 2,  /* not one of the 14 2D ASTC footprints */
    ASTCG_ERR_BUFFER_TOO_SMALL = 3,   /* rgba_len or out_len below what w x h needs */
    ASTCG_ERR_NOT_CONFIGURED = 4,     /* encode before configure */
    ASTCG_ERR_DEVICE = 5,             /* HIP runtime failure, see astcg_last_error */
    ASTCG_ERR_OUT_OF_MEMORY = 6,
    ASTCG_ERR_IO = 7,                 /* file could not be written */
    ASTCG_ERR_UNSUPPORTED = 8         /* e.g. a symbolic block the packer cannot encode */
} astcg_status;

typedef struct astcg_encoder astcg_encoder;

/* Candidate policy.  COMPAT reproduces the reference's search (packed table
 * positions chosen by passes 001-007 including their indexing quirks, all
 * trial modes, 8 refined candidates).  EXHAUSTIVE replaces passes 001-007 by
 * the partition search the reference keeps but never calls
 * (code/best_partitionings.cpp:498-737, caller commented out at
 * code/compression_pipeline.cpp:363-389), run over every selected
 * partitioning; the rest of the search is the same.  LEARNED prunes either
 * with a trained candidate predictor (astcg_set_predictor): no counterpart
 * in the reference. */
typedef enum { ASTCG_POLICY_COMPAT = 0, ASTCG_POLICY_LEARNED = 1, ASTCG_POLICY_EXHAUSTIVE = 2 } astcg_policy;

/* Pruning of the LEARNED policy.  A block whose kept trial modes (its
 * `modes` most likely) hold less than fallback_below of the predicted
 * probability gets the full search (COMPAT, or EXHAUSTIVE with
 * candidates = 1); any other block searches only the
 * partition counts that are its most likely count or
 * at least count_below probable (with the `candidates` search when seeds is
 * 0, else from its `seeds` most likely partitionings of each such count), only its
 * `modes` most likely trial modes, and refines `slots` candidates per
 * partitioned block instead of 8.  With the COMPAT positions, `partitionings`
 * (1..3) keeps only that many of the 4 positions of each searched count: the
 * ones whose seeds the predictor's seed head scores highest.
 * `refine_partitionings` (1..3) prunes later, after the per-mode error
 * estimates: a count's partitionings are ranked by their best estimated
 * candidate error and only the best ones are refined.  `early_out` applies
 * the error limit the reference computes but never uses
 * (code/compression_pipeline.cpp:276-282, 330: 10^(-(80 - 19 log10 T) / 10)
 * * 65535^2 * T * sum of channel weights, T texels per block): once a block's
 * best encoding so far is below early_out times that limit, its remaining
 * partition counts are not searched. */
typedef struct {
    uint32_t seeds;         /* 0 (the `candidates` search, default), 1..4: the predictor's most likely seeds of
                               each searched count, 5..64: those seeds ranked by their line fits (as in the
                               EXHAUSTIVE search) and the 4 best searched */
    uint32_t modes;         /* 1..64, default 16 */
    uint32_t slots;         /* 1..8, default 5 */
    float fallback_below;   /* default 0.5 */
    float count_below;      /* default 0 (every partition count) */
    uint32_t partitionings; /* COMPAT positions kept per count: 1..4, 0 = 4 (default) */
    uint32_t refine_partitionings; /* refine only this many (1..3) of each searched count's partitionings:
                                      those whose best estimated candidate is lowest; 0 = all (default) */
    float early_out;        /* > 0: a block whose best error after a partition count is below early_out
                               times the reference's error limit skips the remaining counts; 0 = off */
    float full_search_above; /* > 0: a block whose best error after one partition is above this multiple
                                of the error limit searches 2..4 partitions with every trial mode and 8
                                slots (the full search); 0 = off */
    uint32_t rounds;        /* refinement rounds per candidate (pass13-17), 1..6, 0 = 6 as the reference */
    uint32_t partitioned_slots; /* candidates refined per partitioned block for 2..4 partitions
                                   (1..8, 0 = slots); fallback blocks keep 8 */
    uint32_t candidates;    /* partition candidates of the searched counts (and of fallback / hard
                               blocks): 0 = the COMPAT positions, 1 = the EXHAUSTIVE search over every
                               selected partitioning, N = 2..1024: that search over the first N in k-means
                               mismatch order (partition_search_limit, code/best_partitionings.cpp:520) */
    float count_above;      /* > 0: a block whose best error so far is above this multiple of the error limit
                               searches every remaining partition count, whatever the count head predicts;
                               0 = off */
    uint32_t full_search_modes; /* hard blocks (full_search_above): 0 = every trial mode and 8 slots (default),
                                   1 = their predicted modes and slots (the full search then widens only the
                                   partition counts and candidates) */
    uint32_t mismatch_candidates; /* seeds 5..64: the line fits first cover this many partitionings of the
                                     k-means mismatch order (as the EXHAUSTIVE search visits them), then the
                                     seeds not among them; 0 = the seeds alone (default) */
    float fallback_above;   /* > 0: a block whose texels lie further from their best line (weighted squared
                               distance) than this multiple of the error limit gets the full search from one
                               partition on, like a low-confidence block; 0 = off */
} astcg_learned_params;

void astcg_default_learned_params(astcg_learned_params* p);

/* Launch shapes of one footprint's search kernels.  No counterpart in the
 * reference (its dispatch sizes are fixed in the WGSL); none of them changes
 * any encoded block, only how the work is laid out on the GPU. */
typedef struct {
    uint32_t refine_stage_max_tc; /* the refinement kernel stages a block's RGBA8 texels in LDS when the footprint has
                                     at most this many texels (default 25) */
    uint32_t ex_f4_max_tc;        /* the exhaustive partition search stages float texels in LDS up to this many
                                     texels, at most 64 (default 16); RGBA8 words above */
    uint32_t exw_max_list;        /* partition searches of block lists (LEARNED: the exhaustive search of hard /
                                     fallback blocks and the seeded search of listed blocks) up to this many
                                     blocks run one wave per block, longer ones one lane per block
                                     (default 65536; 1048576 for footprints of 36 texels or more) */
    uint32_t refine_queue;        /* 0: the refinement runs a workgroup per 64 candidates of one block mode (default);
                                     n >= 1: a persistent work-queue kernel whose waves take those 64-candidate
                                     tiles from a queue and give a lane the next candidate of its tile once n lanes
                                     (at most 64) have finished theirs */
    uint32_t refine_classes;      /* 1: the refinement's candidates are grouped by block mode and endpoint format
                                     class, 0: by block mode only (default 1 where the refinement stages texels) */
} astcg_tuning;

astcg_status astcg_get_tuning(astcg_encoder* enc, uint32_t bx, uint32_t by, astcg_tuning* out);
astcg_status astcg_set_tuning(astcg_encoder* enc, uint32_t bx, uint32_t by, const astcg_tuning* tuning);
/* Chooses the launch shapes of footprint bx x by for this GPU.  With a
 * cache_path that holds an entry for this GPU (architecture and CU count) and
 * footprint, that entry is applied; otherwise each setting is timed on
 * encodes of a synthetic 1024x1024 image (the LDS staging of the refinement
 * and its work queue under the COMPAT search, the float staging of the
 * exhaustive partition search under the EXHAUSTIVE search), the fastest one
 * kept, and, with a cache_path, appended to that file.  An encoder whose footprint tables are
 * first built while the environment variable ASTCG_TUNE_CACHE names such a
 * file applies its entry the same way.  The encoder's configuration is kept
 * (configure is re-run with the last parameters). */
astcg_status astcg_autotune(astcg_encoder* enc, uint32_t bx, uint32_t by, const char* cache_path, astcg_tuning* out);

#define ASTCG_FLAG_PROFILE 1u  /* record HIP events around every kernel launch */
#define ASTCG_FLAG_LABELS 2u   /* keep each partition count's best partitioning (astcg_get_labels) */
/* Blocks whose texels are all equal are written as LDR void-extent blocks
 * (the exact colour); every other block is unchanged.  No counterpart in the
 * reference, which never emits them (code/physical_compression.cpp:849). */
#define ASTCG_FLAG_VOID_EXTENT 4u
/* Capture an encode's whole launch chain into a HIP graph on the first call
 * and replay it on later calls with the same arguments (same buffers, sizes,
 * footprints and settings); up to 16 chains are kept per encoder.  Needs a
 * caller stream (not the null stream); ignored together with
 * ASTCG_FLAG_PROFILE.  No counterpart in the reference. */
#define ASTCG_FLAG_GRAPH 8u
/* After the search, try single-partition dual-plane block modes (one
 * channel on a second weight plane, direct RGB / RGBA endpoints) for every
 * block and keep one where it lowers the block's error.  The reference never
 * tries dual-plane modes (code/compression_pipeline.cpp:183-185); footprints
 * up to 64 texels.  astcg_get_symbolic does not represent the plane-2
 * channel of such blocks. */
#define ASTCG_FLAG_DUAL_PLANE 16u

typedef struct {
    uint32_t max_partitions;  /* 1..4 (reference: 4) */
    uint32_t policy;          /* astcg_policy */
    uint32_t chunk_blocks;    /* blocks per device chunk, 0 = automatic (a working set near 20 GiB) */
    uint32_t flags;           /* ASTCG_FLAG_* */
    uint32_t reserved[4];
} astcg_params;

typedef struct {
    char name[32];
    double total_ms;
    uint64_t launches;
} astcg_kernel_time;

/* SymbolicBlock (code/astc.h:639-655), byte-identical: 432 bytes. */
typedef struct {
    float errorval;
    uint32_t block_mode_index;          /* raw 11-bit block mode */
    uint32_t partition_count;
    uint32_t partition_index;           /* 10-bit partition seed */
    uint32_t partition_formats_matched;
    uint32_t quant_mode;                /* colour endpoint quant level (4..20) */
    uint32_t _padding1;
    uint32_t _padding2;
    uint32_t partition_formats[4];
    uint32_t packed_color_values[32];   /* unquantised 0..255 */
    uint32_t quantized_weights[64];     /* unquantised 0..64 */
} astcg_symbolic_block;

void astcg_default_params(astcg_params* p);

/* ASTCEncoder(device) + init(): binds a HIP device. */
astcg_status astcg_create(int hip_device, astcg_encoder** out);
astcg_status astcg_destroy(astcg_encoder* enc);

/* secondaryInit(w, h, bx, by): per image size and footprint. */
astcg_status astcg_configure(astcg_encoder* enc, uint32_t width, uint32_t height, uint32_t block_x,
                             uint32_t block_y, const astcg_params* params);

/* encode(rgba, out, out_len): synchronous, host buffers.  rgba is row-major
 * RGBA8 (width*height*4 bytes); out receives blocks_x*blocks_y 16-byte
 * blocks in row-major block order. */
astcg_status astcg_encode(astcg_encoder* enc, const uint8_t* rgba, size_t rgba_len, uint8_t* out,
                          size_t out_len);

/* Same on device memory; work is enqueued on `stream` (a hipStream_t, NULL =
 * the null stream) and the call returns without synchronising.  The encoder
 * keeps its scratch buffers until the next configure/destroy. */
astcg_status astcg_encode_device(astcg_encoder* enc, const uint8_t* d_rgba, uint8_t* d_out, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory while they encode (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same encode as
 * astcg_encode_device (identical blocks) that calls no hipMalloc / hipFree
 * or allocating library routine.  Its working memory is one fixed 64 MiB
 * device array per GPU, held until the process ends; the image is searched
 * in chunks that fit it.  astcg_configure (which allocates, once) must have
 * been called; the COMPAT policy only, flags VOID_EXTENT and DUAL_PLANE
 * allowed, PROFILE / LABELS / GRAPH rejected; an explicit chunk_blocks the
 * array cannot hold is rejected with ASTCG_ERR_INVALID_ARGUMENT.  All calls
 * on one GPU share the array and keep its last contents: calls that overlap
 * on one GPU (other streams, other threads, other encoders) race. */
astcg_status astcg_encode_device_noalloc(astcg_encoder* enc, const uint8_t* d_rgba, uint8_t* d_out, void* stream);

/* One texture of a batch.  `rgba` is row-major RGBA8 (width*height*4 bytes),
 * `out` receives ceil(width/block_x)*ceil(height/block_y) 16-byte blocks in
 * row-major block order; out_len is checked against that size. */
typedef struct astcg_texture {
    const uint8_t* rgba;
    uint32_t width, height;
    uint32_t block_x, block_y;
    uint8_t* out;
    size_t out_len;
} astcg_texture;

/* Batched encode of `count` textures of any sizes and footprints.  Each
 * texture's blocks equal what configure(width, height, block_x, block_y) +
 * encode() produce for it alone, i.e. what the reference's
 * ASTCEncoder::secondaryInit + encode (code/compression_pipeline.cpp:263-535)
 * produce per image; textures of one footprint share each kernel launch.
 * Host buffers, synchronous.  `params` may be NULL (defaults); its flags may
 * enable kernel timing.  Does not change the single-image configuration. */
astcg_status astcg_encode_batch(astcg_encoder* enc, const astcg_texture* textures, uint32_t count,
                                const astcg_params* params);

/* Same with `rgba` / `out` in device memory of the encoder's device; work is
 * enqueued on `stream` and the call returns without waiting for it (a call
 * that must grow the encoder's device buffers first waits for the device;
 * calls with the same or smaller batches reuse them).  Successive calls on
 * one encoder must use one stream or order their streams: they share the
 * encoder's buffers. */
astcg_status astcg_encode_batch_device(astcg_encoder* enc, const astcg_texture* textures, uint32_t count,
                                       const astcg_params* params, void* stream);

/* Optional: copy the chosen symbolic blocks of the last encode (host memory,
 * blocks_x*blocks_y entries). */
astcg_status astcg_get_symbolic(astcg_encoder* enc, astcg_symbolic_block* out, size_t count);

/* Installs the candidate predictor of footprint block_x x block_y used by
 * the LEARNED policy: an MLP nin -> hidden -> hidden -> nout with GELU
 * (nin = 3 * block_x * block_y, nout = 4 + 3 * 1024 + trial modes), given as
 * the fp32 concatenation W1 [nin x hidden], b1 [hidden], W2 [hidden x hidden],
 * b2 [hidden], W3 [hidden x nout], b3 [nout] (row-major, `count` floats), and
 * its pruning parameters (NULL = defaults). */
astcg_status astcg_set_predictor(astcg_encoder* enc, uint32_t block_x, uint32_t block_y, uint32_t hidden,
                                 const float* weights, size_t count, const astcg_learned_params* lp);

/* Fraction of the blocks of the last LEARNED encode that fell back to the
 * full search. */
astcg_status astcg_fallback_rate(astcg_encoder* enc, double* rate);

/* Search outcome per block of the last single-image encode made with
 * ASTCG_FLAG_LABELS: 6 int32 per block = {partition count, partition seed,
 * trial block mode index (0..trial modes - 1), best seed for 2, 3 and 4
 * partitions (-1 = none)}.  Training labels for the candidate predictor; no
 * counterpart in the reference. */
astcg_status astcg_get_labels(astcg_encoder* enc, int32_t* out, size_t count);

/* The 4 compat positions of every partition count of the same encode: per
 * block 3 x 4 entries (count 2..4, position 0..3) of the position's
 * partition seed and the best error its search reached (-1 / 1e30 where the
 * position was not searched).  Labels for ranking the positions; no
 * counterpart in the reference. */
astcg_status astcg_get_candidates(astcg_encoder* enc, int16_t* seeds, float* errors, size_t count);

/* Number of trial block modes (the reference's valid_block_modes) of a
 * footprint, and the raw 11-bit mode of trial mode i. */
astcg_status astcg_trial_modes(uint32_t block_x, uint32_t block_y, uint32_t* count, uint16_t* raw_modes,
                               size_t cap);

/* blocks_x, blocks_y and the number of blocks of the configured image. */
astcg_status astcg_block_counts(const astcg_encoder* enc, uint32_t* blocks_x, uint32_t* blocks_y,
                                uint32_t* num_blocks);

/* symbolic_to_physical (host implementation of the device packer). */
astcg_status astcg_symbolic_to_physical(uint32_t block_x, uint32_t block_y, const astcg_symbolic_block* sb,
                                        uint8_t out[16]);

/* encode_ise: bounded integer sequence encoding of count symbols (each below
 * the quant level's alphabet size) at bit_offset of out. */
astcg_status astcg_encode_ise(uint32_t quant, uint32_t count, const uint8_t* in, uint8_t* out,
                              uint32_t bit_offset);

/* get_ise_sequence_bitcount (1024 for an out-of-range quant level). */
uint32_t astcg_ise_bitcount(uint32_t count, uint32_t quant);

/* store_image: 16-byte .astc header (magic 0x5CA1AB13) followed by the blocks. */
astcg_status astcg_store_image(uint32_t block_x, uint32_t block_y, uint32_t width, uint32_t height,
                               const uint8_t* data, size_t data_len, const char* path);

/* create_astc_file_in_memory: writes header + blocks into out (capacity
 * out_cap); *out_len receives 16 + data_len. */
astcg_status astcg_astc_file_in_memory(uint32_t block_x, uint32_t block_y, uint32_t width, uint32_t height,
                                       const uint8_t* data, size_t data_len, uint8_t* out, size_t out_cap,
                                       size_t* out_len);

/* LDR decoder (ASTC specification), host: decodes num_blocks row-major
 * blocks of a width x height image into RGBA8.  Not part of the reference;
 * used to measure quality. */
astcg_status astcg_decode_image(uint32_t block_x, uint32_t block_y, uint32_t width, uint32_t height,
                                const uint8_t* blocks, size_t blocks_len, uint8_t* rgba, size_t rgba_len);

/* The same decoder on the encoder's GPU (identical output; no counterpart in
 * the reference, which has no decoder).  Device buffers, enqueued on
 * `stream`; error blocks decode to magenta as on the host. */
astcg_status astcg_decode_device(astcg_encoder* enc, uint32_t block_x, uint32_t block_y, uint32_t width,
                                 uint32_t height, const uint8_t* d_blocks, uint8_t* d_rgba, void* stream);

/* GPU decode of host buffers (synchronous). */
astcg_status astcg_decode(astcg_encoder* enc, uint32_t block_x, uint32_t block_y, uint32_t width, uint32_t height,
                          const uint8_t* blocks, size_t blocks_len, uint8_t* rgba, size_t rgba_len);

/* Decodes one block to its symbolic form (inverse of symbolic_to_physical for
 * the block modes the encoder emits).  Returns ASTCG_ERR_UNSUPPORTED for
 * void-extent, dual-plane or HDR blocks. */
astcg_status astcg_physical_to_symbolic(uint32_t block_x, uint32_t block_y, const uint8_t in[16],
                                        astcg_symbolic_block* out);

/* Measurement hook with no counterpart in the reference (it has no timers):
 * with ASTCG_FLAG_PROFILE set at configure, per-kernel device time summed
 * over the launches since the last reset.  *count receives the number of
 * kernels with launches; up to cap entries are written. */
astcg_status astcg_kernel_times(astcg_encoder* enc, astcg_kernel_time* out, size_t cap, size_t* count);
astcg_status astcg_reset_kernel_times(astcg_encoder* enc);

const char* astcg_last_error(const astcg_encoder* enc);
const char* astcg_status_string(astcg_status s);

#ifdef __cplusplus
}
#endif
#endif
