/* Internal structures of the CPU oracle (test infrastructure only). */
#ifndef OR_INTERNAL_H
#define OR_INTERNAL_H

#include <math.h>
#include <stdint.h>
#include <string.h>
#include "astc_oracle.h"

#define OR_BAD 0xFFFFu
#define OR_ERR_DEFAULT 1e37f

/* Quant method enum order, astc.h:85-108. */
enum { Q2 = 0, Q3, Q4, Q5, Q6, Q8, Q10, Q12, Q16, Q20, Q24, Q32, Q40, Q48, Q64, Q80, Q96, Q128, Q160, Q192, Q256 };

/* Endpoint formats, pass09_compute_color_error.wgsl:31-46. */
enum {
    FMT_LUMINANCE = 0, FMT_LUMINANCE_DELTA = 1, FMT_LUMINANCE_ALPHA = 4, FMT_LUMINANCE_ALPHA_DELTA = 5,
    FMT_RGB_SCALE = 6, FMT_RGB = 8, FMT_RGB_DELTA = 9, FMT_RGB_SCALE_ALPHA = 10, FMT_RGBA = 12,
    FMT_RGBA_DELTA = 13
};

typedef struct {
    int texel_count, weight_count, wx, wy;
    int max_quant_level, max_angular_steps, max_quant_steps;
    int maxprec_1plane, refprec_1plane;
    int t2w_off;  /* offset into t2w arrays, texel-major */
    int w2t_off;  /* offset into w2t arrays, weight-major */
    uint8_t texel_weight_count[OR_MAX_TEXELS];
    uint16_t texel_weights_offset[OR_MAX_TEXELS];
    uint8_t weight_texel_count[OR_MAX_WEIGHTS];
    uint16_t weight_texels_offset[OR_MAX_WEIGHTS];
} or_dm;

typedef struct {
    int mode_index, decimation_mode, quant_mode, weight_bits, is_dual_plane;
} or_bm;

typedef struct {
    int partition_count, partition_index;
    int texel_count[4];
    uint8_t partition_of_texel[OR_MAX_TEXELS];
} or_pinfo;

typedef struct {
    int xdim, ydim, texel_count;
    int dm_count, bm_count;
    or_dm dm[OR_MAX_DM];
    or_bm bm[OR_MAX_BM];
    uint16_t block_mode_index[OR_MAX_BM];
    /* packed texel->weight (index, contribution) and weight->texel maps */
    int t2w_count, w2t_count;
    uint8_t* t2w_idx; float* t2w_contrib;
    uint8_t* w2t_idx; float* w2t_contrib;
    int ndm, nbm;                 /* valid trial modes */
    int valid_dm[OR_MAX_DM];
    int valid_bm[OR_MAX_BM];      /* packed bm index */
    int valid_bm_dml[OR_MAX_BM];  /* index into valid_dm */
    uint32_t kmeans_texels[64];
    or_pinfo part[3 * OR_PARTITIONINGS + 1];
    int count_selected[4], count_all[4];
    uint64_t coverage[3][OR_PARTITIONINGS][4];
    float sin_table[1024 * 16];
    float cos_table[1024 * 16];
} or_desc;

/* ---------- tables (or_tables.c) ---------- */
extern const int or_levels[21];
void or_btq(int q, int* bits, int* trits, int* quints);
int or_weight_unquant(int q, int code);      /* 0..64 */
int or_color_unquant(int q, int code);       /* 0..255 */
extern int or_quant_to_unquant[12][32];      /* ranks of weight levels */
extern int or_weight_scramble[12][32];       /* rank -> ISE code */
extern uint8_t or_color_u2q[17][512];        /* pass14 COLOR_UNQUANT_TO_QUANT_TABLE */
extern uint8_t or_color_pquant[17][256];     /* uquant -> scrambled code */
extern int8_t or_quant_mode_table[10][128];  /* pass11 QUANT_MODE_TABLE */
extern uint16_t or_prev_next[12][65];        /* pass16 ALL_PREV_NEXT_VALUES */
extern uint8_t or_int_of_trits[243];
extern uint8_t or_int_of_quints[125];
void or_init_tables(void);

or_desc* or_desc_get(int bx, int by);  /* cached per footprint */

/* ---------- WGSL operation semantics ---------- */
static inline float wmin(float a, float b) { return fminf(a, b); }
static inline float wmax(float a, float b) { return fmaxf(a, b); }
static inline float wclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float wround(float x) { return rintf(x); }
static inline int32_t wf2i(float v) {
    if (!(v < 2147483520.0f)) return 2147483647;
    if (v < -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)v;
}
static inline uint32_t wf2u(float v) {
    if (!(v < 4294967040.0f)) return 0xFFFFFFFFu;
    if (v < 0.0f) return 0u;
    return (uint32_t)v;
}
static inline float dot4(const float* a, const float* b) {
    float s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    s = s + a[3] * b[3];
    return s;
}
static inline float dot3(const float* a, const float* b) {
    float s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}
static inline void normalize4(const float* v, float* o) {
    float l = sqrtf(dot4(v, v));
    o[0] = v[0] / l; o[1] = v[1] / l; o[2] = v[2] / l; o[3] = v[3] / l;
}
static inline void normalize3(const float* v, float* o) {
    float l = sqrtf(dot3(v, v));
    o[0] = v[0] / l; o[1] = v[1] / l; o[2] = v[2] / l;
}
float or_atan2f(float y, float x);

/* ---------- packer (or_pack.c) ---------- */
void or_write_bits(unsigned value, unsigned bitcount, unsigned bitoffset, uint8_t* ptr);

#endif
