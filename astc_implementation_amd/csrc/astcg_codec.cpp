// Host bitstream code: bounded integer sequence encoding, symbolic_to_physical,
// the .astc container, and an ASTC LDR decoder used to measure quality.
//
// symbolic_to_physical follows code/physical_compression.cpp:843-1003;
// encode_ise follows physical_compression.cpp:672-826; the container follows
// code/astc_store.cpp:5-104.  The decoder implements the Khronos ASTC LDR
// profile (2D, no HDR endpoint modes) and is not part of the reference.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "astcg_codec.h"
#include "astcg_tables.h"

namespace astcg {

namespace {

void write_bits(unsigned value, unsigned bitcount, unsigned bitoffset, uint8_t* ptr) {
    unsigned mask = (1u << bitcount) - 1u;
    value &= mask;
    ptr += bitoffset >> 3;
    bitoffset &= 7;
    value <<= bitoffset;
    mask = ~(mask << bitoffset);
    ptr[0] = (uint8_t)((ptr[0] & mask) | value);
    ptr[1] = (uint8_t)((ptr[1] & (mask >> 8)) | (value >> 8));
}

unsigned read_bits(const uint8_t* p, int offset, int count) {
    unsigned v = 0;
    for (int i = 0; i < count; ++i) {
        const int b = offset + i;
        v |= (unsigned)((p[b >> 3] >> (b & 7)) & 1) << i;
    }
    return v;
}

int bitrev8(int p) {
    p = ((p & 0x0F) << 4) | ((p >> 4) & 0x0F);
    p = ((p & 0x33) << 2) | ((p >> 2) & 0x33);
    p = ((p & 0x55) << 1) | ((p >> 1) & 0x55);
    return p;
}

}  // namespace

void encode_ise_host(int quant, int count, const uint8_t* in, uint8_t* out, int bit_offset) {
    const QuantTables& qt = quant_tables();
    int bits, trits, quints;
    quant_btq(quant, &bits, &trits, &quints);
    const unsigned mask = (1u << bits) - 1u;
    unsigned off = (unsigned)bit_offset;
    if (trits) {
        static const int tb[5] = {2, 2, 1, 2, 1}, ts[5] = {0, 2, 4, 5, 7};
        for (int i = 0; i < count; i += 5) {
            int t[5] = {0, 0, 0, 0, 0};
            for (int k = 0; k < 5 && i + k < count; ++k) t[k] = in[i + k] >> bits;
            const int T = qt.int_of_trits[(((t[4] * 3 + t[3]) * 3 + t[2]) * 3 + t[1]) * 3 + t[0]];
            for (int k = 0; k < 5 && i + k < count; ++k) {
                const unsigned pack = (in[i + k] & mask) | (((unsigned)(T >> ts[k]) & ((1u << tb[k]) - 1u)) << bits);
                write_bits(pack, bits + tb[k], off, out);
                off += bits + tb[k];
            }
        }
    } else if (quints) {
        static const int qb[3] = {3, 2, 2}, qs[3] = {0, 3, 5};
        for (int i = 0; i < count; i += 3) {
            int t[3] = {0, 0, 0};
            for (int k = 0; k < 3 && i + k < count; ++k) t[k] = in[i + k] >> bits;
            const int Q = qt.int_of_quints[(t[2] * 5 + t[1]) * 5 + t[0]];
            for (int k = 0; k < 3 && i + k < count; ++k) {
                const unsigned pack = (in[i + k] & mask) | (((unsigned)(Q >> qs[k]) & ((1u << qb[k]) - 1u)) << bits);
                write_bits(pack, bits + qb[k], off, out);
                off += bits + qb[k];
            }
        }
    } else {
        for (int i = 0; i < count; ++i) {
            write_bits(in[i], bits, off, out);
            off += bits;
        }
    }
}

// Decodes count ISE symbols starting at bit_offset (ASTC spec integer
// sequence decode).
void decode_ise_host(int quant, int count, const uint8_t* in, int bit_offset, uint8_t* out) {
    int bits, trits, quints;
    quant_btq(quant, &bits, &trits, &quints);
    int off = bit_offset;
    if (trits) {
        static const int tb[5] = {2, 2, 1, 2, 1}, ts[5] = {0, 2, 4, 5, 7};
        for (int i = 0; i < count; i += 5) {
            int m[5] = {0, 0, 0, 0, 0}, T = 0;
            for (int k = 0; k < 5 && i + k < count; ++k) {
                m[k] = (int)read_bits(in, off, bits);
                off += bits;
                T |= (int)read_bits(in, off, tb[k]) << ts[k];
                off += tb[k];
            }
            // spec trit decode of the 8-bit block (all 256 codes, including
            // the ones an encoder never emits)
            const int dg = quant_tables().trit_digits[T];
            for (int k = 0; k < 5 && i + k < count; ++k) out[i + k] = (uint8_t)((((dg >> (2 * k)) & 3) << bits) | m[k]);
        }
    } else if (quints) {
        static const int qb[3] = {3, 2, 2}, qs[3] = {0, 3, 5};
        for (int i = 0; i < count; i += 3) {
            int m[3] = {0, 0, 0}, Q = 0;
            for (int k = 0; k < 3 && i + k < count; ++k) {
                m[k] = (int)read_bits(in, off, bits);
                off += bits;
                Q |= (int)read_bits(in, off, qb[k]) << qs[k];
                off += qb[k];
            }
            const int dg = quant_tables().quint_digits[Q];
            for (int k = 0; k < 3 && i + k < count; ++k) out[i + k] = (uint8_t)((((dg >> (3 * k)) & 7) << bits) | m[k]);
        }
    } else {
        for (int i = 0; i < count; ++i) {
            out[i] = (uint8_t)read_bits(in, off, bits);
            off += bits;
        }
    }
}

int symbolic_to_physical_host(int bx, int by, const astcg_symbolic_block& sb, uint8_t out[16]) {
    const Footprint* f = footprint(bx, by);
    if (!f) return ASTCG_ERR_INVALID_FOOTPRINT;
    if (sb.block_mode_index >= 2048 || f->raw_wc[sb.block_mode_index] == 0) return ASTCG_ERR_UNSUPPORTED;
    const int pc = (int)sb.partition_count;
    if (pc < 1 || pc > 4 || sb.quant_mode < 4 || sb.quant_mode > 20) return ASTCG_ERR_UNSUPPORTED;
    const QuantTables& qt = quant_tables();
    const int raw = (int)sb.block_mode_index;
    const int wc = f->raw_wc[raw], q = f->raw_quant[raw];
    if (f->raw_to_bm[raw] < 0) return ASTCG_ERR_UNSUPPORTED;  // dual-plane modes are not emitted
    const int wbits = ise_bitcount(wc, q);
    static const int kLevels[12] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32};
    uint8_t wbuf[17] = {0}, weights[64];
    const float levels = (float)kLevels[q];
    for (int i = 0; i < wc; ++i) {
        const float uqw = (float)sb.quantized_weights[i];
        const int qwi = (int)((uqw / 64.0f) * (levels - 1.0f) + 0.5f);
        if (qwi < 0 || qwi >= kLevels[q]) return ASTCG_ERR_UNSUPPORTED;
        weights[i] = qt.weight_scramble[q][qwi];
    }
    encode_ise_host(q, wc, weights, wbuf, 0);
    uint8_t blk[17] = {0};
    for (int i = 0; i < 16; ++i) blk[i] = (uint8_t)bitrev8(wbuf[15 - i]);
    write_bits((unsigned)raw, 11, 0, blk);
    write_bits((unsigned)(pc - 1), 2, 11, blk);
    if (pc > 1) {
        write_bits(sb.partition_index & 0x3F, 6, 13, blk);
        write_bits((sb.partition_index >> 6) & 0xF, 4, 19, blk);
        if (sb.partition_formats_matched) {
            write_bits(sb.partition_formats[0] << 2, 6, 23, blk);
        } else {
            int low_class = 4;
            for (int i = 0; i < pc; ++i) low_class = std::min(low_class, (int)(sb.partition_formats[i] >> 2));
            if (low_class == 3) low_class = 2;
            for (int i = 0; i < pc; ++i) {
                const int d = (int)(sb.partition_formats[i] >> 2) - low_class;
                if (d < 0 || d > 1) return ASTCG_ERR_UNSUPPORTED;  // formats span more than two classes
            }
            int enc = low_class + 1, bitpos = 2;
            for (int i = 0; i < pc; ++i) enc |= ((int)(sb.partition_formats[i] >> 2) - low_class) << bitpos++;
            for (int i = 0; i < pc; ++i) { enc |= (int)(sb.partition_formats[i] & 3) << bitpos; bitpos += 2; }
            const int hsize = 3 * pc - 4;
            write_bits((unsigned)(enc & 0x3F), 6, 23, blk);
            write_bits((unsigned)(enc >> 6), (unsigned)hsize, (unsigned)(128 - wbits - hsize), blk);
        }
    } else {
        write_bits(sb.partition_formats[0] & 0xFF, 4, 13, blk);
    }
    uint8_t vals[32];
    int nv = 0;
    for (int i = 0; i < pc; ++i) {
        const int n = 2 * (int)(sb.partition_formats[i] >> 2) + 2;
        if (nv + n > 32) return ASTCG_ERR_UNSUPPORTED;
        for (int j = 0; j < n; ++j) vals[nv + j] = qt.color_pquant[sb.quant_mode - 4][sb.packed_color_values[i * 8 + j] & 0xFF];
        nv += n;
    }
    // the colour integers must end below the weights (and the extra CEM bits)
    const int cem_bits = (pc > 1 && !sb.partition_formats_matched) ? 3 * pc - 4 : 0;
    if ((pc == 1 ? 17 : 29) + ise_bitcount(nv, (int)sb.quant_mode) > 128 - wbits - cem_bits) return ASTCG_ERR_UNSUPPORTED;
    encode_ise_host((int)sb.quant_mode, nv, vals, blk, pc == 1 ? 17 : 29);
    std::memcpy(out, blk, 16);
    return ASTCG_OK;
}

// ---------------------------------------------------------------------------
// Decoder

namespace {

// LDR colour endpoint decode (ASTC spec, endpoint modes 0,1,4,5,6,8,9,10,12,13).
bool decode_endpoints(int cem, const int* v, int e0[4], int e1[4]) {
    auto bts = [](int& a, int& b) {  // bit_transfer_signed(a, b)
        b >>= 1;
        b |= a & 0x80;
        a >>= 1;
        a &= 0x3F;
        if (a & 0x20) a -= 0x40;
    };
    auto blue = [](int* c) {
        c[0] = (c[0] + c[2]) >> 1;
        c[1] = (c[1] + c[2]) >> 1;
    };
    auto clamp255 = [](int x) { return x < 0 ? 0 : (x > 255 ? 255 : x); };
    switch (cem) {
        case 0:
            e0[0] = e0[1] = e0[2] = v[0]; e0[3] = 255;
            e1[0] = e1[1] = e1[2] = v[1]; e1[3] = 255;
            return true;
        case 1: {
            const int l0 = (v[0] >> 2) | (v[1] & 0xC0);
            const int l1 = std::min(l0 + (v[1] & 0x3F), 255);
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
            bts(v1, v0);
            bts(v3, v2);
            e0[0] = e0[1] = e0[2] = v0; e0[3] = v2;
            const int l1 = clamp255(v0 + v1), a1 = clamp255(v2 + v3);
            e1[0] = e1[1] = e1[2] = l1; e1[3] = a1;
            return true;
        }
        case 6:
            for (int c = 0; c < 3; ++c) { e1[c] = v[c]; e0[c] = (v[c] * v[3]) >> 8; }
            e0[3] = e1[3] = 255;
            return true;
        case 10:
            for (int c = 0; c < 3; ++c) { e1[c] = v[c]; e0[c] = (v[c] * v[3]) >> 8; }
            e0[3] = v[4];
            e1[3] = v[5];
            return true;
        case 8:
        case 12: {
            int c0[4] = {v[0], v[2], v[4], cem == 12 ? v[6] : 255};
            int c1[4] = {v[1], v[3], v[5], cem == 12 ? v[7] : 255};
            if (c1[0] + c1[1] + c1[2] < c0[0] + c0[1] + c0[2]) {
                blue(c1);
                blue(c0);
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
            for (int c = 0; c < (cem == 13 ? 4 : 3); ++c) bts(b[c], a[c]);
            int c0[4], c1[4];
            for (int c = 0; c < 4; ++c) { c0[c] = a[c]; c1[c] = a[c] + b[c]; }
            if (cem == 9) { c0[3] = 255; c1[3] = 255; }
            if (b[0] + b[1] + b[2] < 0) {
                blue(c1);
                blue(c0);
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

struct DecodedBlock {
    bool error = false;
    bool void_extent = false;
    int rgba_const[4];
    int pc = 1, seed = 0;
    int cem[4] = {0, 0, 0, 0};
    int color_quant = 0, nvals = 0;
    uint8_t color_codes[32];
    int raw_mode = 0, wq = 0, wx = 0, wy = 0, dual = 0, ccs = 0;
    uint8_t weight_codes[64];
};

bool decode_mode_fields(int mode, int* xw, int* yw, int* dual, int* quant) {
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

DecodedBlock parse_block(int bx, int by, const uint8_t* b) {
    DecodedBlock d;
    const int mode = (int)read_bits(b, 0, 11);
    if ((mode & 0x1FF) == 0x1FC) {
        d.void_extent = true;
        if (!(mode & 0x200)) {  // LDR void extent: 16-bit UNORM colour in bits 64..127
            for (int c = 0; c < 4; ++c) d.rgba_const[c] = (int)read_bits(b, 64 + 16 * c, 16);
        } else {
            d.error = true;
        }
        return d;
    }
    int xw, yw, dual, q;
    if (!decode_mode_fields(mode, &xw, &yw, &dual, &q) || xw > bx || yw > by) { d.error = true; return d; }
    const int wc = xw * yw * (dual + 1);
    const int wbits = ise_bitcount(wc, q);
    if (wc > 64 || wbits < 24 || wbits > 96) { d.error = true; return d; }
    d.raw_mode = mode; d.wq = q; d.wx = xw; d.wy = yw; d.dual = dual;
    d.pc = (int)read_bits(b, 11, 2) + 1;
    if (dual && d.pc == 4) { d.error = true; return d; }
    int below = 128 - wbits;
    int color_start;
    if (d.pc == 1) {
        d.cem[0] = (int)read_bits(b, 13, 4);
        color_start = 17;
    } else {
        d.seed = (int)read_bits(b, 13, 10);
        color_start = 29;
        const int cemlow = (int)read_bits(b, 23, 6);
        if ((cemlow & 3) == 0) {
            for (int i = 0; i < d.pc; ++i) d.cem[i] = cemlow >> 2;
        } else {
            const int hsize = 3 * d.pc - 4;
            below -= hsize;
            const int enc = cemlow | ((int)read_bits(b, below, hsize) << 6);
            const int base_class = (enc & 3) - 1;
            int bitpos = 2;
            int cls[4];
            for (int i = 0; i < d.pc; ++i) cls[i] = base_class + ((enc >> bitpos++) & 1);
            for (int i = 0; i < d.pc; ++i) { d.cem[i] = (cls[i] << 2) | ((enc >> bitpos) & 3); bitpos += 2; }
        }
    }
    if (dual) {
        below -= 2;
        d.ccs = (int)read_bits(b, below, 2);
    }
    d.nvals = 0;
    for (int i = 0; i < d.pc; ++i) d.nvals += 2 * ((d.cem[i] >> 2) + 1);
    if (d.nvals > 18) { d.error = true; return d; }
    const int avail = below - color_start;
    int cq = -1;
    for (int qq = 4; qq <= 20; ++qq)
        if (ise_bitcount(d.nvals, qq) <= avail) cq = qq;
    if (cq < 0) { d.error = true; return d; }
    d.color_quant = cq;
    decode_ise_host(cq, d.nvals, b, color_start, d.color_codes);
    uint8_t rev[16];
    for (int i = 0; i < 16; ++i) rev[i] = (uint8_t)bitrev8(b[15 - i]);
    decode_ise_host(q, wc, rev, 0, d.weight_codes);
    return d;
}

}  // namespace

int physical_to_symbolic_host(int bx, int by, const uint8_t in[16], astcg_symbolic_block* sb) {
    if (!footprint(bx, by)) return ASTCG_ERR_INVALID_FOOTPRINT;
    DecodedBlock d = parse_block(bx, by, in);
    if (d.error || d.void_extent || d.dual) return ASTCG_ERR_UNSUPPORTED;
    std::memset(sb, 0, sizeof(*sb));
    const QuantTables& qt = quant_tables();
    sb->block_mode_index = (uint32_t)d.raw_mode;
    sb->partition_count = (uint32_t)d.pc;
    sb->partition_index = (uint32_t)d.seed;
    bool same = true;
    for (int i = 1; i < d.pc; ++i) same = same && d.cem[i] == d.cem[0];
    sb->partition_formats_matched = (d.pc > 1 && (read_bits(in, 23, 2) == 0)) ? 1u : 0u;
    (void)same;
    sb->quant_mode = (uint32_t)d.color_quant;
    int vi = 0;
    for (int i = 0; i < d.pc; ++i) {
        sb->partition_formats[i] = (uint32_t)d.cem[i];
        const int n = 2 * ((d.cem[i] >> 2) + 1);
        for (int j = 0; j < n; ++j) sb->packed_color_values[i * 8 + j] = (uint32_t)color_unquant(d.color_quant, d.color_codes[vi++]);
    }
    const int wc = d.wx * d.wy;
    for (int i = 0; i < wc; ++i) sb->quantized_weights[i] = (uint32_t)weight_unquant(d.wq, d.weight_codes[i]);
    (void)qt;
    return ASTCG_OK;
}

int decode_block_rgba8(int bx, int by, const uint8_t in[16], uint8_t* texels /* bx*by*4 */) {
    const int tc = bx * by;
    DecodedBlock d = parse_block(bx, by, in);
    if (d.error) {
        for (int t = 0; t < tc; ++t) { texels[t * 4] = 255; texels[t * 4 + 1] = 0; texels[t * 4 + 2] = 255; texels[t * 4 + 3] = 255; }
        return ASTCG_ERR_UNSUPPORTED;
    }
    if (d.void_extent) {
        for (int t = 0; t < tc; ++t)
            for (int c = 0; c < 4; ++c) texels[t * 4 + c] = (uint8_t)((d.rgba_const[c] + 128) / 257);
        return ASTCG_OK;
    }
    // unquantised colour values and endpoints per partition
    int vals[32];
    for (int i = 0; i < d.nvals; ++i) vals[i] = color_unquant(d.color_quant, d.color_codes[i]);
    int e0[4][4], e1[4][4];
    int vi = 0;
    for (int p = 0; p < d.pc; ++p) {
        if (!decode_endpoints(d.cem[p], vals + vi, e0[p], e1[p])) {
            for (int t = 0; t < tc; ++t) { texels[t * 4] = 255; texels[t * 4 + 1] = 0; texels[t * 4 + 2] = 255; texels[t * 4 + 3] = 255; }
            return ASTCG_ERR_UNSUPPORTED;
        }
        vi += 2 * ((d.cem[p] >> 2) + 1);
    }
    const int wc = d.wx * d.wy * (d.dual + 1);
    int uw[64];
    for (int i = 0; i < wc; ++i) uw[i] = weight_unquant(d.wq, d.weight_codes[i]);
    const bool small_block = tc < 31;
    const int plane_stride = d.dual ? 2 : 1;
    for (int y = 0; y < by; ++y)
        for (int x = 0; x < bx; ++x) {
            const int t = y * bx + x;
            // spec weight infill
            const int xwp = (((1024 + bx / 2) / (bx - 1)) * x * (d.wx - 1) + 32) >> 6;
            const int ywp = (((1024 + by / 2) / (by - 1)) * y * (d.wy - 1) + 32) >> 6;
            const int xf = xwp & 0xF, yf = ywp & 0xF, base = (xwp >> 4) + (ywp >> 4) * d.wx;
            const int w11 = (xf * yf + 8) >> 4, w01 = xf - w11, w10 = yf - w11, w00 = 16 - xf - yf + w11;
            int w[2] = {0, 0};
            for (int plane = 0; plane < plane_stride; ++plane) {
                auto W = [&](int idx) { return idx < d.wx * d.wy ? uw[idx * plane_stride + plane] : 0; };
                int s = 0;
                if (w00) s += W(base) * w00;
                if (w01) s += W(base + 1) * w01;
                if (w10) s += W(base + d.wx) * w10;
                if (w11) s += W(base + d.wx + 1) * w11;
                w[plane] = (s + 8) >> 4;
            }
            int p = 0;
            if (d.pc > 1) {
                // spec partition function (same as the encoder's tables)
                const Footprint* f = footprint(bx, by);
                const int e = f->seed_to_packed[(d.pc - 2) * 1024 + d.seed];
                p = f->part_of_texel[((size_t)(d.pc - 2) * 1024 + e) * tc + t];
                (void)small_block;
            }
            for (int c = 0; c < 4; ++c) {
                const int wt = (d.dual && c == d.ccs) ? w[1] : w[0];
                const int c0 = e0[p][c] * 257, c1 = e1[p][c] * 257;
                const int v16 = (c0 * (64 - wt) + c1 * wt + 32) >> 6;
                texels[t * 4 + c] = (uint8_t)((v16 * 255 + 32767) / 65535);
            }
        }
    return ASTCG_OK;
}

}  // namespace astcg
