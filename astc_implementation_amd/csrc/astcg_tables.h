// Host-side precompute for the gfx950 ASTC encoder: spec-derived quantisation
// tables and per-footprint metadata (block modes, decimation grids, partition
// tables, angular sin/cos tables).  Everything here is built once per
// footprint and uploaded to HBM as read-only tables.
#pragma once
#include <cstdint>
#include <vector>

namespace astcg {

constexpr int kMaxTexels = 144;
constexpr int kMaxWeights = 64;
constexpr int kPartitionings = 1024;
constexpr int kRefBatch = 920;  // reference batch geometry (astc.h:739)

// Footprint-independent quantisation tables (ASTC spec formulas).
struct QuantTables {
    uint8_t quant_to_unquant[12][32];  // weight level rank -> 0..64
    uint8_t weight_scramble[12][32];   // weight level rank -> ISE code
    uint16_t prev_next[12][65];        // uquant weight -> (prev | next << 8)
    uint8_t color_u2q[17][512];        // index (2v + r) -> nearest colour level, ties low
    uint8_t color_pquant[17][256];     // colour uquant -> ISE code, ties to larger code
    int8_t quant_mode[10][128];        // (ints / 2, bits) -> best colour quant or -1
    uint8_t int_of_trits[243];
    uint8_t int_of_quints[125];
    float pixel_scale[256];            // (u8 / 255) * 65536
    // decoder side
    uint16_t trit_digits[256];         // 8-bit trit block -> 5 digits, 2 bits each
    uint16_t quint_digits[128];        // 7-bit quint block -> 3 digits, 3 bits each
    uint8_t color_unq[17][256];        // colour ISE code -> 0..255 (quant 4..20)
    uint8_t weight_unq[12][32];        // weight ISE code -> 0..64
};
const QuantTables& quant_tables();

int ise_bitcount(int count, int quant);
int weight_unquant(int quant, int code);
int color_unquant(int quant, int code);
void quant_btq(int quant, int* bits, int* trits, int* quints);

// Per-valid-decimation-mode record (indexed by the trial index).
struct DmInfo {
    uint8_t texel_count, weight_count, wx, wy;
    uint8_t max_quant_level, max_angular_steps, max_quant_steps, pad;
    uint32_t w2t_base;  // first entry of this mode in the weight->texel pool
};

// Per-valid-block-mode record (indexed by the trial index).
struct BmInfo {
    uint16_t raw_mode;   // 11-bit block mode
    uint8_t dm;          // valid decimation mode index
    uint8_t quant;       // weight quant level
    uint8_t weight_bits;
    uint8_t pad[3];
};

// Dual-plane block mode tried by the optional dual-plane search (one
// partition): the decimation grid must be a trial decimation mode.
struct DualMode {
    uint16_t raw_mode;   // 11-bit block mode (dual-plane bit set)
    uint8_t dm;          // trial decimation mode index of its weight grid
    uint8_t quant;       // weight quant level
    int8_t cq_rgb;       // colour quant level left for CEM 8 (6 values), -1 = none
    int8_t cq_rgba;      // colour quant level left for CEM 12 (8 values), -1 = none
    uint8_t pad[2];
};

struct Footprint {
    int bx = 0, by = 0, texel_count = 0;
    int ndm = 0, nbm = 0;
    std::vector<DmInfo> dm;
    std::vector<BmInfo> bm;
    // texel -> weight map, 4 fixed slots per (dm, texel): index and contribution
    std::vector<uint8_t> t2w_count;   // [ndm][texel_count]
    std::vector<uint8_t> t2w_idx;     // [ndm][texel_count][4]
    std::vector<float> t2w_w;         // [ndm][texel_count][4]
    // weight -> texel map: per (dm, weight) count + offset into the pool
    std::vector<uint8_t> w2t_count;   // [ndm][64]
    std::vector<uint16_t> w2t_off;    // [ndm][64], relative to DmInfo::w2t_base
    std::vector<uint8_t> w2t_idx;     // pool
    std::vector<float> w2t_w;         // pool
    // raw block mode -> valid block mode trial (-1 if not encodable here)
    std::vector<int16_t> raw_to_bm;   // [2048]
    // raw block mode -> (weight count, quant, weight bits, dual plane) for the packer;
    // wc == 0 marks modes invalid for this footprint
    std::vector<uint8_t> raw_wc, raw_quant, raw_wbits;
    std::vector<DualMode> dual;       // dual-plane modes of the optional dual-plane search
    // partition tables, packed order, for partition counts 2..4 (+ the 1-partition entry)
    std::vector<uint8_t> part_of_texel;  // [3][1024][texel_count]
    std::vector<uint16_t> part_seed;     // [3][1024]
    std::vector<uint8_t> part_counts;    // [3][1024][4]
    int count_selected[4] = {1, 0, 0, 0};
    std::vector<uint64_t> coverage;      // [3][1024][4]
    uint32_t kmeans_texels[64] = {};
    // partial-bitonic orderings (first 128 entries) for score0 == 0 and > 0, per count 2..4
    uint16_t ordering[3][2][128] = {};
    std::vector<float> sin_table, cos_table;  // [1024][16]
    // seed -> packed table index, per partition count 2..4 (0xFFFF if absent)
    std::vector<uint16_t> seed_to_packed;     // [3][1024]
};

// Builds (and caches) the metadata for one of the 14 2D footprints.
// Returns nullptr for footprints ASTC does not define.
const Footprint* footprint(int bx, int by);
bool valid_footprint(int bx, int by);

}  // namespace astcg
