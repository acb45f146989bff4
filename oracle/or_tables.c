/*
 * Oracle tables and per-footprint metadata (test infrastructure only).
 *
 * Every lookup table the reference hard-codes is regenerated here from the
 * ASTC specification formulas:
 *   - weight unquantisation / QUANT_TO_UNQUANT (pass07_weights_and_error_for_bm.wgsl:12-37)
 *   - weight scramble maps (physical_compression.cpp:249-274)
 *   - colour unquantisation, COLOR_UNQUANT_TO_QUANT_TABLE (pass14_pack_color_endpoints.wgsl:29)
 *     = nearest level to index/2, ties to the lower level
 *   - color_uquant_to_scrambled_pquant_tables (physical_compression.cpp:284-591)
 *     = nearest level, ties to the larger ISE code
 *   - QUANT_MODE_TABLE (pass11_best_color_combination_for_mode_1part.wgsl:12-63)
 *   - ALL_PREV_NEXT_VALUES (pass16_realign_weights.wgsl:13)
 *   - integer_of_trits / integer_of_quints (physical_compression.cpp:44-243)
 *     = inverse of the spec trit/quint decode, largest code among duplicates
 * Metadata follows metadata_structures.cpp and partition_tables.cpp.
 */
#include <stdlib.h>
#include <stdio.h>
#include <pthread.h>
#include "or_internal.h"

const int or_levels[21] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256};

int or_quant_to_unquant[12][32];
int or_weight_scramble[12][32];
uint8_t or_color_u2q[17][512];
uint8_t or_color_pquant[17][256];
int8_t or_quant_mode_table[10][128];
uint16_t or_prev_next[12][65];
uint8_t or_int_of_trits[243];
uint8_t or_int_of_quints[125];

void or_btq(int q, int* bits, int* trits, int* quints) {
    int n = or_levels[q];
    int t = 0, u = 0, base = n;
    if (n % 3 == 0) { t = 1; base = n / 3; }
    else if (n % 5 == 0) { u = 1; base = n / 5; }
    int b = 0;
    while ((1 << b) < base) b++;
    *bits = b; *trits = t; *quints = u;
}

/* physical_compression.cpp:625-639 (ceil(bits*n + 8n/5 or 7n/3)) */
int oracle_ise_bitcount(int count, int quant) {
    if (quant < 0 || quant >= 21) return 1024;
    int b, t, u;
    or_btq(quant, &b, &t, &u);
    return b * count + (t ? (8 * count + 4) / 5 : 0) + (u ? (7 * count + 2) / 3 : 0);
}

static int replicate(int v, int bits, int to) {
    int r = 0, sh = to;
    while (sh > 0) {
        if (sh >= bits) { r |= v << (sh - bits); sh -= bits; }
        else { r |= v >> (bits - sh); sh = 0; }
    }
    return r;
}

/* ASTC weight unquantisation (spec), result 0..64. */
int or_weight_unquant(int q, int code) {
    int b, t, u;
    or_btq(q, &b, &t, &u);
    int T;
    if (!t && !u) {
        T = replicate(code, b, 6);
    } else {
        int D = code >> b, m = code & ((1 << b) - 1);
        if (b == 0) {
            static const int t3[3] = {0, 32, 63}, t5[5] = {0, 16, 32, 47, 63};
            T = t ? t3[D] : t5[D];
        } else {
            int a = m & 1, bb = (m >> 1) & 1, c = (m >> 2) & 1;
            int A = a ? 0x7F : 0, B = 0, C;
            if (t) {
                C = (b == 1) ? 50 : (b == 2) ? 23 : 11;
                if (b == 2) B = (bb << 6) | (bb << 2) | bb;
                if (b == 3) B = (c << 6) | (bb << 5) | (c << 1) | bb;
            } else {
                C = (b == 1) ? 28 : 13;
                if (b == 2) B = (bb << 6) | (bb << 1);
            }
            T = D * C + B;
            T ^= A;
            T = (A & 0x20) | (T >> 2);
        }
    }
    if (T > 32) T += 1;
    return T;
}

/* ASTC colour endpoint unquantisation (spec), result 0..255. */
int or_color_unquant(int q, int code) {
    int b, t, u;
    or_btq(q, &b, &t, &u);
    if (!t && !u) return replicate(code, b, 8);
    int D = code >> b, m = code & ((1 << b) - 1);
    int bit[6] = {0};
    for (int i = 0; i < b; i++) bit[i] = (m >> i) & 1;
    int a = bit[0], bb = bit[1], c = bit[2], d = bit[3], e = bit[4], f = bit[5];
    int A = a ? 0x1FF : 0, B = 0, C = 0;
    if (t) {
        switch (b) {
        case 1: C = 204; B = 0; break;
        case 2: C = 93; B = (bb << 8) | (bb << 4) | (bb << 2) | (bb << 1); break;
        case 3: C = 44; B = (c << 8) | (bb << 7) | (c << 3) | (bb << 2) | (c << 1) | bb; break;
        case 4: C = 22; B = (d << 8) | (c << 7) | (bb << 6) | (d << 2) | (c << 1) | bb; break;
        case 5: C = 11; B = (e << 8) | (d << 7) | (c << 6) | (bb << 5) | (e << 1) | d; break;
        case 6: C = 5;  B = (f << 8) | (e << 7) | (d << 6) | (c << 5) | (bb << 4) | f; break;
        }
    } else {
        switch (b) {
        case 1: C = 113; B = 0; break;
        case 2: C = 54; B = (bb << 8) | (bb << 3) | (bb << 2); break;
        case 3: C = 26; B = (c << 8) | (bb << 7) | (c << 2) | (bb << 1) | c; break;
        case 4: C = 13; B = (d << 8) | (c << 7) | (bb << 6) | (d << 1) | c; break;
        case 5: C = 6;  B = (e << 8) | (d << 7) | (c << 6) | (bb << 5) | e; break;
        }
    }
    int T = D * C + B;
    T ^= A;
    T = (A & 0x80) | (T >> 2);
    return T;
}

/* ASTC spec trit decode: 8-bit T -> five trits. */
static void trit_decode(int T, int* t) {
    int C;
    if (((T >> 2) & 7) == 7) {
        C = (((T >> 5) & 7) << 2) | (T & 3);
        t[4] = 2; t[3] = 2;
    } else {
        C = T & 0x1F;
        if (((T >> 5) & 3) == 3) { t[4] = 2; t[3] = (T >> 7) & 1; }
        else { t[4] = (T >> 7) & 1; t[3] = (T >> 5) & 3; }
    }
    if ((C & 3) == 3) {
        t[2] = 2; t[1] = (C >> 4) & 1;
        t[0] = (((C >> 3) & 1) << 1) | (((C >> 2) & 1) & ~((C >> 3) & 1) & 1);
    } else if (((C >> 2) & 3) == 3) {
        t[2] = 2; t[1] = 2; t[0] = C & 3;
    } else {
        t[2] = (C >> 4) & 1; t[1] = (C >> 2) & 3;
        t[0] = (((C >> 1) & 1) << 1) | ((C & 1) & ~((C >> 1) & 1) & 1);
    }
}

/* ASTC spec quint decode: 7-bit Q -> three quints. */
static void quint_decode(int Q, int* q) {
    if (((Q >> 1) & 3) == 3 && ((Q >> 5) & 3) == 0) {
        int q0b = Q & 1;
        q[2] = (q0b << 2) | ((((Q >> 4) & 1) & ~q0b & 1) << 1) | (((Q >> 3) & 1) & ~q0b & 1);
        q[1] = 4; q[0] = 4;
    } else {
        int C;
        if (((Q >> 1) & 3) == 3) {
            q[2] = 4;
            C = (((Q >> 3) & 3) << 3) | (((~(Q >> 5)) & 3) << 1) | (Q & 1);
        } else {
            q[2] = (Q >> 5) & 3;
            C = Q & 0x1F;
        }
        if ((C & 7) == 5) { q[1] = 4; q[0] = (C >> 3) & 3; }
        else { q[1] = (C >> 3) & 3; q[0] = C & 7; }
    }
}

static pthread_once_t tables_once = PTHREAD_ONCE_INIT;

static void build_tables(void) {
    for (int q = 0; q < 12; q++) {
        int n = or_levels[q];
        int vals[32], codes[32];
        for (int c = 0; c < n; c++) { vals[c] = or_weight_unquant(q, c); codes[c] = c; }
        /* sort codes by unquantised value (values are distinct) */
        for (int i = 1; i < n; i++)
            for (int j = i; j > 0 && vals[codes[j]] < vals[codes[j - 1]]; j--) {
                int tmp = codes[j]; codes[j] = codes[j - 1]; codes[j - 1] = tmp;
            }
        for (int r = 0; r < n; r++) {
            or_quant_to_unquant[q][r] = vals[codes[r]];
            or_weight_scramble[q][r] = codes[r];
        }
        for (int v = 0; v < 65; v++) or_prev_next[q][v] = 0;
        for (int r = 0; r < n; r++) {
            int prv = or_quant_to_unquant[q][r > 0 ? r - 1 : 0];
            int nxt = or_quant_to_unquant[q][r < n - 1 ? r + 1 : n - 1];
            or_prev_next[q][or_quant_to_unquant[q][r]] = (uint16_t)(prv | (nxt << 8));
        }
    }
    for (int q = 4; q < 21; q++) {
        int n = or_levels[q];
        int lv[256];
        for (int c = 0; c < n; c++) lv[c] = or_color_unquant(q, c);
        for (int i = 0; i < 512; i++) {
            /* nearest level to i/2, ties to the lower level */
            int best = -1, bd = 1 << 30;
            for (int c = 0; c < n; c++) {
                int d = abs(2 * lv[c] - i);
                if (d < bd || (d == bd && lv[c] < lv[best])) { bd = d; best = c; }
            }
            or_color_u2q[q - 4][i] = (uint8_t)lv[best];
        }
        for (int u = 0; u < 256; u++) {
            /* nearest level to u, ties to the larger ISE code */
            int best = -1, bd = 1 << 30;
            for (int c = 0; c < n; c++) {
                int d = abs(lv[c] - u);
                if (d < bd || (d == bd && c > best)) { bd = d; best = c; }
            }
            or_color_pquant[q - 4][u] = (uint8_t)best;
        }
    }
    for (int row = 0; row < 10; row++) {
        for (int bits = 0; bits < 128; bits++) {
            int best = -1;
            if (row > 0)
                for (int q = 0; q < 21; q++)
                    if (oracle_ise_bitcount(2 * row, q) <= bits) best = q;
            or_quant_mode_table[row][bits] = (int8_t)best;
        }
    }
    for (int T = 0; T < 256; T++) {
        int t[5];
        trit_decode(T, t);
        int idx = (((t[4] * 3 + t[3]) * 3 + t[2]) * 3 + t[1]) * 3 + t[0];
        or_int_of_trits[idx] = (uint8_t)T; /* ascending T: the largest duplicate wins */
    }
    for (int Q = 0; Q < 128; Q++) {
        int q[3];
        quint_decode(Q, q);
        int idx = (q[2] * 5 + q[1]) * 5 + q[0];
        or_int_of_quints[idx] = (uint8_t)Q;
    }
}

void or_init_tables(void) { pthread_once(&tables_once, build_tables); }

/* Table ids: 0 QUANT_TO_UNQUANT (flat, 142), 1 weight scramble (flat, 142),
 * 2 COLOR_UNQUANT_TO_QUANT (8704), 3 pquant (4352), 4 QUANT_MODE_TABLE (1280),
 * 5 ALL_PREV_NEXT (780), 6 integer_of_trits (243), 7 integer_of_quints (125) */
int oracle_table(int id, int32_t* out, int cap) {
    or_init_tables();
    int n = 0;
#define PUT(v) do { if (out && n < cap) out[n] = (int32_t)(v); n++; } while (0)
    switch (id) {
    case 0: for (int q = 0; q < 12; q++) for (int r = 0; r < or_levels[q]; r++) PUT(or_quant_to_unquant[q][r]); break;
    case 1: for (int q = 0; q < 12; q++) for (int r = 0; r < or_levels[q]; r++) PUT(or_weight_scramble[q][r]); break;
    case 2: for (int q = 0; q < 17; q++) for (int i = 0; i < 512; i++) PUT(or_color_u2q[q][i]); break;
    case 3: for (int q = 0; q < 17; q++) for (int i = 0; i < 256; i++) PUT(or_color_pquant[q][i]); break;
    case 4: for (int r = 0; r < 10; r++) for (int b = 0; b < 128; b++) PUT(or_quant_mode_table[r][b]); break;
    case 5: for (int q = 0; q < 12; q++) for (int v = 0; v < 65; v++) PUT(or_prev_next[q][v]); break;
    case 6: for (int i = 0; i < 243; i++) PUT(or_int_of_trits[i]); break;
    case 7: for (int i = 0; i < 125; i++) PUT(or_int_of_quints[i]); break;
    default: return -1;
    }
#undef PUT
    return n;
}

/* ------------------------------------------------------------------ */
/* Per-footprint metadata: metadata_structures.cpp, partition_tables.cpp */

/* decode_block_mode_2d, metadata_structures.cpp:72-173 */
static int decode_block_mode_2d(int mode, int* xw, int* yw, int* dual, int* quant, int* wbits) {
    int base_q = (mode >> 4) & 1, H = (mode >> 9) & 1, D = (mode >> 10) & 1, A = (mode >> 5) & 3;
    *xw = 0; *yw = 0;
    if ((mode & 3) != 0) {
        base_q |= (mode & 3) << 1;
        int B = (mode >> 7) & 3;
        switch ((mode >> 2) & 3) {
        case 0: *xw = B + 4; *yw = A + 2; break;
        case 1: *xw = B + 8; *yw = A + 2; break;
        case 2: *xw = A + 2; *yw = B + 8; break;
        case 3:
            B &= 1;
            if (mode & 0x100) { *xw = B + 2; *yw = A + 2; }
            else { *xw = A + 2; *yw = B + 6; }
            break;
        }
    } else {
        base_q |= ((mode >> 2) & 3) << 1;
        if (((mode >> 2) & 3) == 0) return 0;
        int B = (mode >> 9) & 3;
        switch ((mode >> 7) & 3) {
        case 0: *xw = 12; *yw = A + 2; break;
        case 1: *xw = A + 2; *yw = 12; break;
        case 2: *xw = A + 6; *yw = B + 6; D = 0; H = 0; break;
        case 3:
            switch ((mode >> 5) & 3) {
            case 0: *xw = 6; *yw = 10; break;
            case 1: *xw = 10; *yw = 6; break;
            default: return 0;
            }
            break;
        }
    }
    int wc = (*xw) * (*yw) * (D + 1);
    *quant = (base_q - 2) + 6 * H;
    *dual = D != 0;
    *wbits = oracle_ise_bitcount(wc, *quant);
    return wc <= 64 && *wbits >= 24 && *wbits <= 96;
}

static const int steps_for_quant_level[12] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32};

/* init_decimation_info + construct_dt_entry, metadata_structures.cpp:176-336 */
static void build_dm(or_desc* d, int idx, int xw, int yw) {
    or_dm* m = &d->dm[idx];
    int xt = d->xdim, yt = d->ydim, tc = xt * yt, wc = xw * yw;
    uint8_t wcount_of_texel[OR_MAX_TEXELS], grid_w[OR_MAX_TEXELS][4], w_of_texel[OR_MAX_TEXELS][4];
    uint8_t tcount_of_weight[OR_MAX_WEIGHTS];
    static __thread uint8_t texels_of_w[OR_MAX_WEIGHTS][OR_MAX_TEXELS], tw_of_w[OR_MAX_WEIGHTS][OR_MAX_TEXELS];
    memset(wcount_of_texel, 0, sizeof(wcount_of_texel));
    memset(tcount_of_weight, 0, sizeof(tcount_of_weight));
    for (int y = 0; y < yt; y++)
        for (int x = 0; x < xt; x++) {
            int texel = y * xt + x;
            int x_weight = (((1024 + xt / 2) / (xt - 1)) * x * (xw - 1) + 32) >> 6;
            int y_weight = (((1024 + yt / 2) / (yt - 1)) * y * (yw - 1) + 32) >> 6;
            int xf = x_weight & 0xF, yf = y_weight & 0xF, xi = x_weight >> 4, yi = y_weight >> 4;
            int qw[4], w[4];
            qw[0] = xi + yi * xw; qw[1] = qw[0] + 1; qw[2] = qw[0] + xw; qw[3] = qw[2] + 1;
            int prod = xf * yf;
            w[3] = (prod + 8) >> 4;
            w[1] = xf - w[3];
            w[2] = yf - w[3];
            w[0] = 16 - xf - yf + w[3];
            for (int i = 0; i < 4; i++)
                if (w[i] != 0) {
                    int ti = wcount_of_texel[texel]++;
                    grid_w[texel][ti] = (uint8_t)qw[i];
                    w_of_texel[texel][ti] = (uint8_t)w[i];
                    int wi = tcount_of_weight[qw[i]]++;
                    texels_of_w[qw[i]][wi] = (uint8_t)texel;
                    tw_of_w[qw[i]][wi] = (uint8_t)w[i];
                }
        }
    m->t2w_off = d->t2w_count;
    for (int i = 0; i < tc; i++) {
        m->texel_weight_count[i] = wcount_of_texel[i];
        m->texel_weights_offset[i] = (uint16_t)(i == 0 ? 0 : m->texel_weights_offset[i - 1] + m->texel_weight_count[i - 1]);
        for (int j = 0; j < wcount_of_texel[i]; j++) {
            d->t2w_idx[d->t2w_count] = grid_w[i][j];
            d->t2w_contrib[d->t2w_count] = (float)w_of_texel[i][j] * (1.0f / 16.0f);
            d->t2w_count++;
        }
    }
    m->w2t_off = d->w2t_count;
    for (int i = 0; i < wc; i++) {
        m->weight_texel_count[i] = tcount_of_weight[i];
        m->weight_texels_offset[i] = (uint16_t)(i == 0 ? 0 : m->weight_texels_offset[i - 1] + m->weight_texel_count[i - 1]);
        for (int j = 0; j < tcount_of_weight[i]; j++) {
            d->w2t_idx[d->w2t_count] = texels_of_w[i][j];
            d->w2t_contrib[d->w2t_count] = (float)tw_of_w[i][j] * (1.0f / 16.0f);
            d->w2t_count++;
        }
    }
    m->texel_count = tc; m->weight_count = wc; m->wx = xw; m->wy = yw;
    int maxprec = -1;
    for (int q = 0; q < 12; q++) {
        int bits = oracle_ise_bitcount(wc, q);
        if (bits >= 24 && bits <= 96) maxprec = q;
    }
    m->maxprec_1plane = maxprec;
    m->refprec_1plane = 0;
    /* quant_limit (astc.h:143) = QUANT_12, TUNE_MAX_ANGULAR_QUANT = QUANT_12 */
    int mp = maxprec < 7 ? maxprec : 7;
    if (mp < 0) mp = 0; /* unreachable for dms with 1-plane modes */
    m->max_quant_level = mp;
    m->max_angular_steps = steps_for_quant_level[mp];
    m->max_quant_steps = steps_for_quant_level[mp];
}

/* xorshift RNG, metadata_structures.cpp:3-23 */
static uint64_t rotl64(uint64_t v, int c) { return (v << c) | (v >> (64 - c)); }

/* hash52 and select_partition, partition_tables.cpp:92-241 */
static uint32_t hash52(uint32_t inp) {
    inp ^= inp >> 15;
    inp *= 0xEEDE0891u;
    inp ^= inp >> 5;
    inp += inp << 16;
    inp ^= inp >> 7;
    inp ^= inp >> 3;
    inp ^= inp << 6;
    inp ^= inp >> 17;
    return inp;
}

static int select_partition(int seed, int x, int y, int z, int pc, int small) {
    if (small) { x <<= 1; y <<= 1; z <<= 1; }
    seed += (pc - 1) * 1024;
    uint32_t r = hash52((uint32_t)seed);
    uint8_t s[13];
    s[1] = r & 0xF; s[2] = (r >> 4) & 0xF; s[3] = (r >> 8) & 0xF; s[4] = (r >> 12) & 0xF;
    s[5] = (r >> 16) & 0xF; s[6] = (r >> 20) & 0xF; s[7] = (r >> 24) & 0xF; s[8] = (r >> 28) & 0xF;
    s[9] = (r >> 18) & 0xF; s[10] = (r >> 22) & 0xF; s[11] = (r >> 26) & 0xF;
    s[12] = ((r >> 30) | (r << 2)) & 0xF;
    for (int i = 1; i <= 12; i++) s[i] = (uint8_t)(s[i] * s[i]);
    int sh1, sh2;
    if (seed & 1) { sh1 = (seed & 2) ? 4 : 5; sh2 = (pc == 3) ? 6 : 5; }
    else { sh1 = (pc == 3) ? 6 : 5; sh2 = (seed & 2) ? 4 : 5; }
    int sh3 = (seed & 0x10) ? sh1 : sh2;
    s[1] >>= sh1; s[2] >>= sh2; s[3] >>= sh1; s[4] >>= sh2;
    s[5] >>= sh1; s[6] >>= sh2; s[7] >>= sh1; s[8] >>= sh2;
    s[9] >>= sh3; s[10] >>= sh3; s[11] >>= sh3; s[12] >>= sh3;
    int a = s[1] * x + s[2] * y + s[11] * z + (int)(r >> 14);
    int b = s[3] * x + s[4] * y + s[12] * z + (int)(r >> 10);
    int c = s[5] * x + s[6] * y + s[9] * z + (int)(r >> 6);
    int d = s[7] * x + s[8] * y + s[10] * z + (int)(r >> 2);
    a &= 0x3F; b &= 0x3F; c &= 0x3F; d &= 0x3F;
    if (pc <= 3) d = 0;
    if (pc <= 2) c = 0;
    if (pc <= 1) b = 0;
    if (a >= b && a >= c && a >= d) return 0;
    if (b >= c && b >= d) return 1;
    if (c >= d) return 2;
    return 3;
}

/* generate_one_partition_info_entry, partition_tables.cpp:243-335 */
static int gen_pinfo(or_desc* d, int pc, int seed, int remap, or_pinfo* pi) {
    int small = d->texel_count < 32;
    int counts[4] = {0, 0, 0, 0};
    int t = 0;
    for (int y = 0; y < d->ydim; y++)
        for (int x = 0; x < d->xdim; x++) {
            int p = select_partition(seed, x, y, 0, pc, small);
            counts[p]++;
            pi->partition_of_texel[t++] = (uint8_t)p;
        }
    if (counts[0] == 0) pi->partition_count = 0;
    else if (counts[1] == 0) pi->partition_count = 1;
    else if (counts[2] == 0) pi->partition_count = 2;
    else if (counts[3] == 0) pi->partition_count = 3;
    else pi->partition_count = 4;
    pi->partition_index = seed;
    for (int i = 0; i < 4; i++) pi->texel_count[i] = counts[i];
    if (pc >= 2) {
        uint64_t* bm = d->coverage[pc - 2][remap];
        for (int i = 0; i < pc; i++) bm[i] = 0;
        int n = d->texel_count < 64 ? d->texel_count : 64;
        for (int i = 0; i < n; i++) bm[pi->partition_of_texel[d->kmeans_texels[i]]] |= 1ull << i;
    }
    return pi->partition_count == pc;
}

/* generate_canonical_partitioning, partition_tables.cpp:16-49 */
static void canonical(int tc, const uint8_t* pot, uint64_t* bits) {
    int map[4] = {-1, -1, -1, -1}, n = 0;
    for (int i = 0; i < 5; i++) bits[i] = 0;
    for (int i = 0; i < tc; i++) {
        int p = pot[i];
        if (map[p] < 0) map[p] = n++;
        bits[i >> 5] |= (uint64_t)map[p] << (2 * (i & 0x1F));
    }
}

/* build_partition_table_for_one_partition_count, partition_tables.cpp:337-413 */
static void build_ptab(or_desc* d, int pc) {
    or_pinfo* tab = d->part + (pc - 2) * OR_PARTITIONINGS;
    static __thread uint64_t canon[OR_PARTITIONINGS][5];
    uint8_t built[OR_PARTITIONINGS];
    memset(built, 0, sizeof(built));
    int next = 0;
    d->count_selected[pc - 1] = 0;
    d->count_all[pc - 1] = 0;
    for (int pass = 0; pass < 2; pass++) {
        for (int i = 0; i < OR_PARTITIONINGS; i++) {
            if (pass == 1 && built[i]) continue;
            int useful = gen_pinfo(d, pc, i, next, &tab[next]);
            if (pass == 0 && !useful) continue;
            canonical(d->texel_count, tab[next].partition_of_texel, canon[next]);
            int keep = 1;
            for (int j = 0; j < next; j++)
                if (memcmp(canon[next], canon[j], sizeof(canon[0])) == 0) { keep = 0; break; }
            if (useful && keep) {
                if (pass == 0) {
                    d->count_selected[pc - 1]++;
                    d->count_all[pc - 1]++;
                    built[i] = 1;
                    next++;
                }
            } else if (pass == 1) {
                d->count_all[pc - 1]++;
                next++;
            }
        }
    }
}

static void build_desc(or_desc* d, int bx, int by) {
    memset(d, 0, sizeof(*d));
    d->xdim = bx; d->ydim = by; d->texel_count = bx * by;
    d->t2w_idx = (uint8_t*)malloc(OR_MAX_DM * OR_MAX_TEXELS * 4);
    d->t2w_contrib = (float*)malloc(sizeof(float) * OR_MAX_DM * OR_MAX_TEXELS * 4);
    d->w2t_idx = (uint8_t*)malloc(OR_MAX_DM * OR_MAX_TEXELS * 4);
    d->w2t_contrib = (float*)malloc(sizeof(float) * OR_MAX_DM * OR_MAX_TEXELS * 4);
    int dmi[12 * 16 + 12];
    for (int i = 0; i < 12 * 16 + 12; i++) dmi[i] = -1;
    int nbm = 0, ndm = 0;
    /* construct_metadata_structures, metadata_structures.cpp:338-461 */
    for (int i = 0; i < OR_MAX_BM; i++) {
        int xw, yw, dual, quant, wbits;
        int valid = decode_block_mode_2d(i, &xw, &yw, &dual, &quant, &wbits);
        if (!valid || xw > bx || yw > by) { d->block_mode_index[i] = OR_BAD; continue; }
        if (dual ? (109 - wbits) <= 0 : (111 - wbits) <= 0) { d->block_mode_index[i] = OR_BAD; continue; }
        int dm = dmi[yw * 16 + xw];
        if (dm < 0) { build_dm(d, ndm, xw, yw); dm = ndm; dmi[yw * 16 + xw] = ndm; ndm++; }
        or_bm* b = &d->bm[nbm];
        b->decimation_mode = dm; b->quant_mode = quant; b->is_dual_plane = dual;
        b->weight_bits = wbits; b->mode_index = i;
        if (!dual) d->dm[dm].refprec_1plane |= 1 << quant;
        d->block_mode_index[i] = (uint16_t)nbm;
        nbm++;
    }
    d->bm_count = nbm; d->dm_count = ndm;
    /* assign_kmeans_texels, metadata_structures.cpp:33-70.  Entries past the
     * texel count stay 0 (the reference reads 64 entries; its heap-allocated
     * descriptor is zero-filled there). */
    if (d->texel_count <= 64) {
        for (int i = 0; i < d->texel_count; i++) d->kmeans_texels[i] = (uint32_t)i;
    } else {
        uint64_t st[2] = {0xfaf9e171cea1ec6bULL, 0xf1b318cc06af5d71ULL};
        uint8_t seen[OR_MAX_TEXELS];
        memset(seen, 0, sizeof(seen));
        int set = 0;
        while (set < 64) {
            uint64_t s0 = st[0], s1 = st[1], res = s0 + s1;
            s1 ^= s0;
            st[0] = rotl64(s0, 24) ^ s1 ^ (s1 << 16);
            st[1] = rotl64(s1, 37);
            uint8_t texel = (uint8_t)res;
            texel = (uint8_t)(texel % d->texel_count);
            if (!seen[texel]) { d->kmeans_texels[set++] = texel; seen[texel] = 1; }
        }
    }
    /* init_partition_tables, partition_tables.cpp:415-436 */
    gen_pinfo(d, 1, 0, 0, &d->part[3 * OR_PARTITIONINGS]);
    d->count_selected[0] = 1; d->count_all[0] = 1;
    build_ptab(d, 2);
    build_ptab(d, 3);
    build_ptab(d, 4);
    /* initTrialModes, compression_pipeline.cpp:146-198: weight quant <= QUANT_12 */
    int remap[OR_MAX_DM];
    d->ndm = 0;
    for (int m = 0; m < ndm; m++) {
        remap[m] = -1;
        if ((d->dm[m].refprec_1plane & 0xFF) == 0) continue;
        remap[m] = d->ndm;
        d->valid_dm[d->ndm++] = m;
    }
    d->nbm = 0;
    for (int i = 0; i < OR_MAX_BM; i++) {
        if (d->block_mode_index[i] == OR_BAD) continue;
        or_bm* b = &d->bm[d->block_mode_index[i]];
        if (b->is_dual_plane) continue;
        if (remap[b->decimation_mode] < 0) continue;
        d->valid_bm[d->nbm] = d->block_mode_index[i];
        d->valid_bm_dml[d->nbm] = remap[b->decimation_mode];
        d->nbm++;
    }
    /* construct_angular_tables, metadata_structures.cpp:463-481 */
    const float PI = 3.14159265358979323846f;
    for (int i = 0; i < 16; i++) {
        float step = (float)(i + 1);
        for (int j = 0; j < 1024; j++) {
            float angle = (2.0f * PI / (1024 - 1.0f)) * step * (float)j;
            d->sin_table[j * 16 + i] = sinf(angle);
            d->cos_table[j * 16 + i] = cosf(angle);
        }
    }
}

static pthread_mutex_t desc_lock = PTHREAD_MUTEX_INITIALIZER;
static or_desc* desc_cache[13][13];

or_desc* or_desc_get(int bx, int by) {
    if (bx < 4 || by < 4 || bx > 12 || by > 12) return NULL;
    or_init_tables();
    pthread_mutex_lock(&desc_lock);
    or_desc* d = desc_cache[bx][by];
    if (!d) {
        d = (or_desc*)malloc(sizeof(or_desc));
        build_desc(d, bx, by);
        desc_cache[bx][by] = d;
    }
    pthread_mutex_unlock(&desc_lock);
    return d;
}

int oracle_footprint_info(int bx, int by, int32_t* counts, int cap) {
    or_desc* d = or_desc_get(bx, by);
    if (!d) return -1;
    int32_t v[8] = {d->dm_count, d->bm_count, d->ndm, d->nbm, d->count_selected[1], d->count_selected[2],
                    d->count_selected[3], d->t2w_count};
    for (int i = 0; i < 8 && i < cap; i++) counts[i] = v[i];
    return 8;
}

int oracle_partition_entry(int bx, int by, int pc, int idx, uint8_t* pot, int32_t* sc) {
    or_desc* d = or_desc_get(bx, by);
    if (!d || pc < 1 || pc > 4 || idx < 0 || idx >= OR_PARTITIONINGS) return -1;
    const or_pinfo* pi = pc == 1 ? &d->part[3 * OR_PARTITIONINGS] : &d->part[(pc - 2) * OR_PARTITIONINGS + idx];
    memcpy(pot, pi->partition_of_texel, (size_t)d->texel_count);
    sc[0] = pi->partition_index;
    for (int i = 0; i < 4; i++) sc[1 + i] = pi->texel_count[i];
    sc[5] = pi->partition_count;
    return d->texel_count;
}

/* Decimation table of trial decimation mode `m`: out = {texel_count,
 * weight_count, wx, wy, then per texel: count, 4 x (weight index, 16*contribution)} */
int oracle_decimation_entry(int bx, int by, int m, int32_t* out, int cap) {
    or_desc* d = or_desc_get(bx, by);
    if (!d || m < 0 || m >= d->ndm) return -1;
    const or_dm* di = &d->dm[d->valid_dm[m]];
    int n = 0;
#define PUT(v) do { if (n < cap) out[n] = (int32_t)(v); n++; } while (0)
    PUT(di->texel_count); PUT(di->weight_count); PUT(di->wx); PUT(di->wy);
    for (int t = 0; t < di->texel_count; t++) {
        int c = di->texel_weight_count[t];
        int off = di->t2w_off + di->texel_weights_offset[t];
        PUT(c);
        for (int j = 0; j < 4; j++) {
            if (j < c) { PUT(d->t2w_idx[off + j]); PUT((int)(d->t2w_contrib[off + j] * 16.0f + 0.5f)); }
            else { PUT(-1); PUT(0); }
        }
    }
#undef PUT
    return n;
}
