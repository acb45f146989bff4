/*
 * Oracle: bounded integer sequence encoding and 128-bit block assembly
 * (test infrastructure only).  Follows physical_compression.cpp:625-1003.
 */
#include "or_internal.h"

/* write_bits, physical_compression.cpp:652-670 */
void or_write_bits(unsigned value, unsigned bitcount, unsigned bitoffset, uint8_t* ptr) {
    unsigned mask = (1u << bitcount) - 1u;
    value &= mask;
    ptr += bitoffset >> 3;
    bitoffset &= 7;
    value <<= bitoffset;
    mask <<= bitoffset;
    mask = ~mask;
    ptr[0] = (uint8_t)((ptr[0] & mask) | value);
    ptr[1] = (uint8_t)((ptr[1] & (mask >> 8)) | (value >> 8));
}

/* encode_ise, physical_compression.cpp:672-826 */
void oracle_encode_ise(int quant, int count, const uint8_t* in, uint8_t* out, int bit_offset) {
    or_init_tables();
    int bits, trits, quints;
    or_btq(quant, &bits, &trits, &quints);
    unsigned mask = (1u << bits) - 1u;
    unsigned off = (unsigned)bit_offset;
    if (trits) {
        static const int tb[5] = {2, 2, 1, 2, 1}, ts[5] = {0, 2, 4, 5, 7};
        int i = 0;
        for (int blk = 0; blk < count / 5; blk++) {
            int t[5];
            for (int k = 0; k < 5; k++) t[k] = in[i + k] >> bits;
            int T = or_int_of_trits[(((t[4] * 3 + t[3]) * 3 + t[2]) * 3 + t[1]) * 3 + t[0]];
            for (int k = 0; k < 5; k++) {
                unsigned pack = (in[i] & mask) | (((unsigned)(T >> ts[k]) & ((1u << tb[k]) - 1u)) << bits);
                or_write_bits(pack, (unsigned)(bits + tb[k]), off, out);
                off += (unsigned)(bits + tb[k]);
                i++;
            }
        }
        if (i != count) {
            int t[5] = {0, 0, 0, 0, 0};
            for (int k = 0; k < 4; k++) t[k] = (i + k < count) ? (in[i + k] >> bits) : 0;
            int T = or_int_of_trits[(((t[4] * 3 + t[3]) * 3 + t[2]) * 3 + t[1]) * 3 + t[0]];
            for (int k = 0; i < count; i++, k++) {
                unsigned pack = (in[i] & mask) | (((unsigned)(T >> ts[k]) & ((1u << tb[k]) - 1u)) << bits);
                or_write_bits(pack, (unsigned)(bits + tb[k]), off, out);
                off += (unsigned)(bits + tb[k]);
            }
        }
    } else if (quints) {
        static const int qb[3] = {3, 2, 2}, qs[3] = {0, 3, 5};
        int i = 0;
        for (int blk = 0; blk < count / 3; blk++) {
            int q[3];
            for (int k = 0; k < 3; k++) q[k] = in[i + k] >> bits;
            int Q = or_int_of_quints[(q[2] * 5 + q[1]) * 5 + q[0]];
            for (int k = 0; k < 3; k++) {
                unsigned pack = (in[i] & mask) | (((unsigned)(Q >> qs[k]) & ((1u << qb[k]) - 1u)) << bits);
                or_write_bits(pack, (unsigned)(bits + qb[k]), off, out);
                off += (unsigned)(bits + qb[k]);
                i++;
            }
        }
        if (i != count) {
            int q[3] = {0, 0, 0};
            for (int k = 0; k < 2; k++) q[k] = (i + k < count) ? (in[i + k] >> bits) : 0;
            int Q = or_int_of_quints[(q[2] * 5 + q[1]) * 5 + q[0]];
            for (int k = 0; i < count; i++, k++) {
                unsigned pack = (in[i] & mask) | (((unsigned)(Q >> qs[k]) & ((1u << qb[k]) - 1u)) << bits);
                or_write_bits(pack, (unsigned)(bits + qb[k]), off, out);
                off += (unsigned)(bits + qb[k]);
            }
        }
    } else {
        for (int i = 0; i < count; i++) {
            or_write_bits(in[i], (unsigned)bits, off, out);
            off += (unsigned)bits;
        }
    }
}

static int bitrev8(int p) {
    p = ((p & 0x0F) << 4) | ((p >> 4) & 0x0F);
    p = ((p & 0x33) << 2) | ((p >> 2) & 0x33);
    p = ((p & 0x55) << 1) | ((p >> 1) & 0x55);
    return p;
}

/* symbolic_to_physical, physical_compression.cpp:843-1003 */
int oracle_symbolic_to_physical(int bx, int by, const or_symbolic* sb, uint8_t out[16]) {
    or_desc* d = or_desc_get(bx, by);
    if (!d || sb->block_mode_index >= 2048 || d->block_mode_index[sb->block_mode_index] == OR_BAD) return -1;
    if (sb->quant_mode < 4 || sb->quant_mode > 20) return -2;
    int pc = (int)sb->partition_count;
    if (pc < 1 || pc > 4) return -3;
    const or_bm* bm = &d->bm[d->block_mode_index[sb->block_mode_index]];
    const or_dm* di = &d->dm[bm->decimation_mode];
    int wc = di->weight_count, q = bm->quant_mode;
    float levels = (float)or_levels[q];
    int real_wc = bm->is_dual_plane ? 2 * wc : wc;
    int wbits = oracle_ise_bitcount(real_wc, q);
    uint8_t wbuf[16] = {0}, weights[64];
    for (int i = 0; i < wc; i++) {
        float uqw = (float)sb->quantized_weights[i];
        float qw = (uqw / 64.0f) * (levels - 1.0f);
        int qwi = (int)(qw + 0.5f);
        weights[i] = (uint8_t)or_weight_scramble[q][qwi];
    }
    oracle_encode_ise(q, real_wc, weights, wbuf, 0);
    for (int i = 0; i < 16; i++) out[i] = (uint8_t)bitrev8(wbuf[15 - i]);
    or_write_bits(sb->block_mode_index & 0xFFFFu, 11, 0, out);
    or_write_bits((unsigned)(pc - 1), 2, 11, out);
    if (pc > 1) {
        or_write_bits(sb->partition_index, 6, 13, out);
        or_write_bits(sb->partition_index >> 6, 4, 19, out);
        if (sb->partition_formats_matched) {
            or_write_bits(sb->partition_formats[0] << 2, 6, 23, out);
        } else {
            int low_class = 4;
            for (int i = 0; i < pc; i++) {
                int cls = (int)(sb->partition_formats[i] >> 2);
                if (cls < low_class) low_class = cls;
            }
            if (low_class == 3) low_class = 2;
            int enc = low_class + 1, bitpos = 2;
            for (int i = 0; i < pc; i++) { enc |= ((int)(sb->partition_formats[i] >> 2) - low_class) << bitpos; bitpos++; }
            for (int i = 0; i < pc; i++) { enc |= (int)(sb->partition_formats[i] & 3) << bitpos; bitpos += 2; }
            int lowpart = enc & 0x3F, highpart = enc >> 6;
            int hsize = 3 * pc - 4;
            int hpos = 128 - wbits - hsize;
            or_write_bits((unsigned)lowpart, 6, 23, out);
            or_write_bits((unsigned)highpart, (unsigned)hsize, (unsigned)hpos, out);
        }
    } else {
        or_write_bits((uint8_t)sb->partition_formats[0], 4, 13, out);
    }
    uint8_t vals[32];
    int nv = 0;
    const uint8_t* tab = or_color_pquant[sb->quant_mode - Q6];
    for (int i = 0; i < pc; i++) {
        int n = 2 * (int)(sb->partition_formats[i] >> 2) + 2;
        for (int j = 0; j < n; j++) vals[nv + j] = tab[sb->packed_color_values[i * 8 + j] & 0xFF];
        nv += n;
    }
    oracle_encode_ise((int)sb->quant_mode, nv, vals, out, pc == 1 ? 17 : 29);
    return 0;
}
