/*
 * Oracle: colour endpoint packing and unpacking (test infrastructure only).
 *
 * pack_color_endpoints_helper and the quantisers follow
 * shaders/pass14_pack_color_endpoints.wgsl:391-1221; the unpackers follow
 * shaders/pass15_unpack_color_endpoints.wgsl:95-305.
 */
#include "or_color.h"

static int u2q(int q, int index) { return or_color_u2q[q - 4][index]; }

/* quant_color3_int, pass14:403-412 */
static void quant3_int(int q, const int* v, int* o) {
    for (int c = 0; c < 3; c++) o[c] = u2q(q, v[c] * 2 + 1);
    o[3] = 0;
}

/* quant_color3_flt, pass14:414-429 */
static void quant3_flt(int q, const int* v, const float* vf, int* o) {
    for (int c = 0; c < 3; c++) {
        int idx = v[c] * 2;
        float residual = vf[c] - (float)v[c];
        if (residual >= -0.1f) idx += 1;
        o[c] = u2q(q, idx);
    }
    o[3] = 0;
}

/* quant_color_int, pass14:431-435 */
static int quant1_int(int q, int v) { return u2q(q, v * 2 + 1); }

/* quant_color_flt, pass14:437-448 */
static int quant1_flt(int q, int v, float vf) {
    int idx = v * 2;
    float residual = vf - (float)v;
    if (residual >= -0.1f) idx += 1;
    return u2q(q, idx);
}

/* bit_transfer_signed, pass14:450-466 (input0, input1 by reference) */
static void bit_transfer_signed(int* in0, int* in1) {
    for (int c = 0; c < 4; c++) {
        uint32_t a = (uint32_t)in0[c], b = (uint32_t)in1[c];
        int nb = (int)((b >> 1) | (a & 0x80u));
        int na = (int)((a >> 1) & 0x3Fu);
        if (na & 0x20) na -= 0x40;
        in0[c] = na;
        in1[c] = nb;
    }
}

/* uncontract_color, pass14:468-472 */
static void uncontract(int* v) {
    int b = v[2];
    int r = (v[0] + b) >> 1, g = (v[1] + b) >> 1;
    v[0] = r; v[1] = g;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* rgba_delta_unpack, pass14:474-500 / pass15:267-292 */
void or_rgba_delta_unpack(const int* in0, const int* in1, int* o0, int* o1) {
    int i0[4], i1[4];
    memcpy(i0, in0, sizeof(i0)); memcpy(i1, in1, sizeof(i1));
    bit_transfer_signed(i1, i0);
    int rgb_sum = i1[0] + i1[1] + i1[2];
    for (int c = 0; c < 4; c++) i1[c] = i1[c] + i0[c];
    if (rgb_sum < 0) {
        uncontract(i0); uncontract(i1);
        int t[4];
        memcpy(t, i0, sizeof(t)); memcpy(i0, i1, sizeof(t)); memcpy(i1, t, sizeof(t));
    }
    for (int c = 0; c < 4; c++) { o0[c] = clampi(i0[c], 0, 255); o1[c] = clampi(i1[c], 0, 255); }
}

/* rgba_unpack, pass14:502-523 / pass15:233-254 */
void or_rgba_unpack(const int* in0, const int* in1, int* o0, int* o1) {
    int i0[4], i1[4];
    memcpy(i0, in0, sizeof(i0)); memcpy(i1, in1, sizeof(i1));
    if ((i0[0] + i0[1] + i0[2]) > (i1[0] + i1[1] + i1[2])) {
        uncontract(i0); uncontract(i1);
        int t[4];
        memcpy(t, i0, sizeof(t)); memcpy(i0, i1, sizeof(t)); memcpy(i1, t, sizeof(t));
    }
    memcpy(o0, i0, sizeof(i0)); memcpy(o1, i1, sizeof(i1));
}

/* get_rgba_encoding_error, pass14:391-401 */
static float enc_error(const float* u0, const float* u1, const int* q0, const int* q1) {
    float e[4];
    for (int c = 0; c < 4; c++) {
        float e0 = u0[c] - (float)q0[c], e1 = u1[c] - (float)q1[c];
        e[c] = e0 * e0 + e1 * e1;
    }
    return ((e[0] + e[1]) + e[2]) + e[3];
}

static int any_out_of_range(const float* c) {
    for (int i = 0; i < 4; i++) if (c[i] < 0.0f || c[i] > 255.0f) return 1;
    return 0;
}

/* try_quantize_rgb_delta_blue_contract, pass14:525-606 */
static int try_rgb_delta_bc(const float* color0, const float* color1, int* out0, int* out1, int q) {
    float c0[4], c1[4];
    for (int i = 0; i < 4; i++) { c0[i] = color1[i]; c1[i] = color0[i]; }
    float b0 = c0[2], b1 = c1[2], a0 = c0[3], a1 = c1[3];
    float s0[4] = {b0, b0, b0, a0}, s1[4] = {b1, b1, b1, a1};
    for (int i = 0; i < 4; i++) { c0[i] = (c0[i] + c0[i]) - s0[i]; c1[i] = (c1[i] + c1[i]) - s1[i]; }
    if (any_out_of_range(c0) || any_out_of_range(c1)) return 0;
    int c0a[4], c0b[4], c0be[4], c1d[4], c1de[4];
    for (int i = 0; i < 4; i++) c0a[i] = wf2i(wround(c0[i])) << 1;
    for (int i = 0; i < 4; i++) c0b[i] = c0a[i] & 0xFF;
    quant3_int(q, c0b, c0be);
    for (int i = 0; i < 4; i++) c0b[i] = c0be[i] | (c0a[i] & 0x100);
    for (int i = 0; i < 4; i++) c1d[i] = wf2i(wround(c1[i])) << 1;
    for (int i = 0; i < 4; i++) c1d[i] = c1d[i] - c0b[i];
    c1d[3] = 0;
    for (int i = 0; i < 4; i++) if (c1d[i] > 63 || c1d[i] < -64) return 0;
    for (int i = 0; i < 4; i++) c1d[i] = (c1d[i] & 0x7F) | ((c0b[i] & 0x100) >> 1);
    quant3_int(q, c1d, c1de);
    for (int i = 0; i < 3; i++) if (((c1d[i] ^ c1de[i]) & 0xC0) != 0) return 0;
    int ep0[4], ep1[4];
    memcpy(ep0, c0be, sizeof(ep0)); memcpy(ep1, c1de, sizeof(ep1));
    bit_transfer_signed(ep1, ep0);
    if ((ep1[0] + ep1[1] + ep1[2]) >= 0) return 0;
    for (int i = 0; i < 4; i++) {
        ep0[i] = ep0[i] + ep1[i];
        if (ep0[i] < 0 || ep0[i] > 0xFF) return 0;
    }
    memcpy(out0, c0be, sizeof(c0be)); memcpy(out1, c1de, sizeof(c1de));
    return 1;
}

/* try_quantize_rgb_delta, pass14:608-672 */
static int try_rgb_delta(const float* color0, const float* color1, int* out0, int* out1, int q) {
    int c0a[4], c0b[4], c0be[4], c1d[4], c1de[4];
    for (int i = 0; i < 4; i++) c0a[i] = wf2i(wround(color0[i])) << 1;
    for (int i = 0; i < 4; i++) c0b[i] = c0a[i] & 0xFF;
    quant3_int(q, c0b, c0be);
    for (int i = 0; i < 4; i++) c0b[i] = c0be[i] | (c0a[i] & 0x100);
    for (int i = 0; i < 4; i++) c1d[i] = (wf2i(wround(color1[i])) << 1) - c0b[i];
    c1d[3] = 0;
    for (int i = 0; i < 4; i++) if (c1d[i] > 63 || c1d[i] < -64) return 0;
    for (int i = 0; i < 4; i++) c1d[i] = (c1d[i] & 0x7F) | ((c0b[i] & 0x100) >> 1);
    quant3_int(q, c1d, c1de);
    for (int i = 0; i < 3; i++) if (((c1d[i] ^ c1de[i]) & 0xC0) != 0) return 0;
    int ep0[4], ep1[4];
    memcpy(ep0, c0be, sizeof(ep0)); memcpy(ep1, c1de, sizeof(ep1));
    bit_transfer_signed(ep1, ep0);
    if ((ep1[0] + ep1[1] + ep1[2]) < 0) return 0;
    for (int i = 0; i < 4; i++) {
        ep0[i] = ep0[i] + ep1[i];
        if (ep0[i] < 0 || ep0[i] > 0xFF) return 0;
    }
    memcpy(out0, c0be, sizeof(c0be)); memcpy(out1, c1de, sizeof(c1de));
    return 1;
}

/* try_quantize_rgb_blue_contract, pass14:674-711 */
static int try_rgb_bc(const float* color0, const float* color1, int* out0, int* out1, int q) {
    float c0[4], c1[4];
    float s0[4] = {color0[2], color0[2], color0[2], color0[3]};
    float s1[4] = {color1[2], color1[2], color1[2], color1[3]};
    for (int i = 0; i < 4; i++) { c0[i] = (color0[i] + color0[i]) - s0[i]; c1[i] = (color1[i] + color1[i]) - s1[i]; }
    if (any_out_of_range(c0) || any_out_of_range(c1)) return 0;
    int r0[4], r1[4], c0i[4], c1i[4];
    for (int i = 0; i < 4; i++) { r0[i] = wf2i(wround(c0[i])); r1[i] = wf2i(wround(c1[i])); }
    quant3_flt(q, r0, c0, c0i);
    quant3_flt(q, r1, c1, c1i);
    if ((c1i[0] + c1i[1] + c1i[2]) <= (c0i[0] + c0i[1] + c0i[2])) return 0;
    memcpy(out0, c1i, sizeof(c1i)); memcpy(out1, c0i, sizeof(c0i));
    return 1;
}

/* quantize_rgb, pass14:715-746 */
static void quantize_rgb(const float* color0, const float* color1, int* out0, int* out1, int q) {
    float c0[4], c1[4];
    memcpy(c0, color0, sizeof(c0)); memcpy(c1, color1, sizeof(c1));
    int c0i[4], c1i[4];
    for (;;) {
        int c0q[4], c1q[4];
        for (int i = 0; i < 4; i++) { int r = wf2i(wround(c0[i])); c0q[i] = r > 0 ? r : 0; }
        quant3_flt(q, c0q, c0, c0i);
        for (int i = 0; i < 4; i++) c0[i] = c0[i] - 0.2f;
        for (int i = 0; i < 4; i++) { int r = wf2i(wround(c1[i])); c1q[i] = r < 255 ? r : 255; }
        quant3_flt(q, c1q, c1, c1i);
        for (int i = 0; i < 4; i++) c1[i] = c1[i] + 0.2f;
        if ((c0i[0] + c0i[1] + c0i[2]) <= (c1i[0] + c1i[1] + c1i[2])) break;
    }
    memcpy(out0, c0i, sizeof(c0i)); memcpy(out1, c1i, sizeof(c1i));
}

/* try_quantize_alpha_delta, pass14:748-797 */
static int try_alpha_delta(const float* color0, const float* color1, int* out0, int* out1, int q) {
    int a0a = wf2i(wround(color0[3])) << 1;
    int a0b = a0a & 0xFF;
    int a0be = quant1_int(q, a0b);
    a0b = a0be | (a0a & 0x100);
    int a1d = (wf2i(wround(color1[3])) << 1) - a0b;
    if (a1d > 63 || a1d < -64) return 0;
    a1d = (a1d & 0x7F) | ((a0b & 0x100) >> 1);
    int a1de = quant1_int(q, a1d);
    int a1du = a1de;
    if (((a1d ^ a1du) & 0xC0) != 0) return 0;
    a1du &= 0x7F;
    if (a1du & 0x40) a1du -= 0x80;
    a1du += a0b;
    if (a1du < 0 || a1du > 0x1FF) return 0;
    out0[3] = a0be;
    out1[3] = a1de;
    return 1;
}

/* quantize_rgbs, pass14:858-884 */
static void quantize_rgbs(const float* color, uint32_t* out, int q) {
    const float scale = 1.0f / 257.0f;
    float r = wclamp(color[0] * scale, 0.0f, 255.0f);
    float g = wclamp(color[1] * scale, 0.0f, 255.0f);
    float b = wclamp(color[2] * scale, 0.0f, 255.0f);
    int ri = quant1_flt(q, wf2i(wround(r)), r);
    int gi = quant1_flt(q, wf2i(wround(g)), g);
    int bi = quant1_flt(q, wf2i(wround(b)), b);
    float oldsum = ((color[0] + color[1]) + color[2]) * scale;
    float newsum = (float)(uint32_t)(ri + gi + bi);
    float scalea = wclamp((color[3] * (oldsum + 1e-10f)) / (newsum + 1e-10f), 0.0f, 1.0f);
    int sidx = wf2i(wround(scalea * 256.0f));
    sidx = clampi(sidx, 0, 255);
    out[0] = (uint32_t)ri; out[1] = (uint32_t)gi; out[2] = (uint32_t)bi;
    out[3] = (uint32_t)quant1_int(q, sidx);
}

static float lum3(const float* c) { return ((c[0] + c[1]) + c[2]) / 3.0f; }

/* quantize_luminance, pass14:902-921 */
static void quantize_luminance(const float* c0, const float* c1, uint32_t* out, int q) {
    float l0 = lum3(c0), l1 = lum3(c1);
    if (l0 > l1) { float avg = (l0 + l1) * 0.5f; l0 = avg; l1 = avg; }
    out[0] = (uint32_t)quant1_flt(q, wf2i(wround(l0)), l0);
    out[1] = (uint32_t)quant1_flt(q, wf2i(wround(l1)), l1);
}

/* try_quantize_luminance_alpha_delta, pass14:923-995 */
static int try_la_delta(const float* c0, const float* c1, uint32_t* out, int q) {
    float l0 = lum3(c0), l1 = lum3(c1), a0 = c0[3], a1 = c1[3];
    int l0a = wf2i(wround(l0)) << 1, a0a = wf2i(wround(a0)) << 1;
    int l0b = l0a & 0xFF, a0b = a0a & 0xFF;
    int l0be = quant1_int(q, l0b), a0be = quant1_int(q, a0b);
    l0b = l0be | (l0a & 0x100);
    a0b = a0be | (a0a & 0x100);
    int l1d = (wf2i(wround(l1)) << 1) - l0b;
    int a1d = (wf2i(wround(a1)) << 1) - a0b;
    if (l1d > 63 || l1d < -64 || a1d > 63 || a1d < -64) return 0;
    l1d = (l1d & 0x7F) | ((l0b & 0x100) >> 1);
    a1d = (a1d & 0x7F) | ((a0b & 0x100) >> 1);
    int l1de = quant1_int(q, l1d), a1de = quant1_int(q, a1d);
    int l1du = l1de, a1du = a1de;
    if (((l1d ^ l1du) & 0xC0) != 0) return 0;
    if (((a1d ^ a1du) & 0xC0) != 0) return 0;
    l1du &= 0x7F; a1du &= 0x7F;
    if (l1du & 0x40) l1du -= 0x80;
    if (a1du & 0x40) a1du -= 0x80;
    l1du += l0b; a1du += a0b;
    if (l1du < 0 || l1du > 0x1FF || a1du < 0 || a1du > 0x1FF) return 0;
    out[0] = (uint32_t)l0be; out[1] = (uint32_t)l1de; out[2] = (uint32_t)a0be; out[3] = (uint32_t)a1de;
    return 1;
}

/* quantize_luminance_alpha, pass14:997-1013 */
static void quantize_la(const float* c0, const float* c1, uint32_t* out, int q) {
    float l0 = lum3(c0), l1 = lum3(c1);
    out[0] = (uint32_t)quant1_flt(q, wf2i(wround(l0)), l0);
    out[1] = (uint32_t)quant1_flt(q, wf2i(wround(l1)), l1);
    out[2] = (uint32_t)quant1_flt(q, wf2i(wround(c0[3])), c0[3]);
    out[3] = (uint32_t)quant1_flt(q, wf2i(wround(c1[3])), c1[3]);
}

/* pack_color_endpoints_helper, pass14:1018-1221 */
int or_pack_color_endpoints(const float* color0, const float* color1, const float* rgbs, int format, int q,
                            uint32_t* packed) {
    float c0[4], c1[4];
    for (int i = 0; i < 4; i++) {
        c0[i] = wclamp(color0[i], 0.0f, 65535.0f) / 257.0f;
        c1[i] = wclamp(color1[i], 0.0f, 65535.0f) / 257.0f;
    }
    float best_error = OR_ERR_DEFAULT;
    int best_format = FMT_LUMINANCE;
    for (int i = 0; i < 8; i++) packed[i] = 0;
    int o0[4] = {0, 0, 0, 0}, o1[4] = {0, 0, 0, 0}, t0[4] = {0, 0, 0, 0}, t1[4] = {0, 0, 0, 0};
    int u0[4], u1[4];
    switch (format) {
    case FMT_RGB:
        if (q <= Q160) {
            if (try_rgb_delta_bc(c0, c1, o0, o1, q)) {
                or_rgba_delta_unpack(o0, o1, u0, u1);
                best_format = FMT_RGB_DELTA;
                best_error = enc_error(c0, c1, u0, u1);
            }
            if (try_rgb_delta(c0, c1, t0, t1, q)) {
                or_rgba_delta_unpack(t0, t1, u0, u1);
                float e = enc_error(c0, c1, u0, u1);
                if (e < best_error) { best_error = e; best_format = FMT_RGB_DELTA; memcpy(o0, t0, sizeof(o0)); memcpy(o1, t1, sizeof(o1)); }
            }
        }
        if (q < Q256) {
            if (try_rgb_bc(c0, c1, t0, t1, q)) {
                or_rgba_unpack(t0, t1, u0, u1);
                float e = enc_error(c0, c1, u0, u1);
                if (e < best_error) { best_error = e; best_format = FMT_RGB; memcpy(o0, t0, sizeof(o0)); memcpy(o1, t1, sizeof(o1)); }
            }
        }
        quantize_rgb(c0, c1, t0, t1, q);
        or_rgba_unpack(t0, t1, u0, u1);
        {
            float e = enc_error(c0, c1, u0, u1);
            if (e < best_error) { best_error = e; best_format = FMT_RGB; memcpy(o0, t0, sizeof(o0)); memcpy(o1, t1, sizeof(o1)); }
        }
        packed[0] = (uint32_t)o0[0]; packed[1] = (uint32_t)o1[0];
        packed[2] = (uint32_t)o0[1]; packed[3] = (uint32_t)o1[1];
        packed[4] = (uint32_t)o0[2]; packed[5] = (uint32_t)o1[2];
        break;
    case FMT_RGBA:
        if (q <= Q160) {
            /* try_quantize_rgba_delta_blue_contract: rgb part then alpha with swapped colours */
            if (try_rgb_delta_bc(c0, c1, o0, o1, q) && try_alpha_delta(c1, c0, o0, o1, q)) {
                or_rgba_delta_unpack(o0, o1, u0, u1);
                best_format = FMT_RGBA_DELTA;
                best_error = enc_error(c0, c1, u0, u1);
            }
            if (try_rgb_delta(c0, c1, t0, t1, q) && try_alpha_delta(c0, c1, t0, t1, q)) {
                or_rgba_delta_unpack(t0, t1, u0, u1);
                float e = enc_error(c0, c1, u0, u1);
                if (e < best_error) { best_error = e; best_format = FMT_RGBA_DELTA; memcpy(o0, t0, sizeof(o0)); memcpy(o1, t1, sizeof(o1)); }
            }
        }
        if (q < Q256) {
            if (try_rgb_bc(c0, c1, t0, t1, q)) {
                t0[3] = quant1_flt(q, wf2i(wround(c1[3])), c1[3]);
                t1[3] = quant1_flt(q, wf2i(wround(c0[3])), c0[3]);
                or_rgba_unpack(t0, t1, u0, u1);
                float e = enc_error(c0, c1, u0, u1);
                if (e < best_error) { best_error = e; best_format = FMT_RGBA; memcpy(o0, t0, sizeof(o0)); memcpy(o1, t1, sizeof(o1)); }
            }
        }
        quantize_rgb(c0, c1, t0, t1, q);
        t0[3] = quant1_flt(q, wf2i(wround(c0[3])), c0[3]);
        t1[3] = quant1_flt(q, wf2i(wround(c1[3])), c1[3]);
        or_rgba_unpack(t0, t1, u0, u1);
        {
            float e = enc_error(c0, c1, u0, u1);
            if (e < best_error) { best_error = e; best_format = FMT_RGBA; memcpy(o0, t0, sizeof(o0)); memcpy(o1, t1, sizeof(o1)); }
        }
        packed[0] = (uint32_t)o0[0]; packed[1] = (uint32_t)o1[0];
        packed[2] = (uint32_t)o0[1]; packed[3] = (uint32_t)o1[1];
        packed[4] = (uint32_t)o0[2]; packed[5] = (uint32_t)o1[2];
        packed[6] = (uint32_t)o0[3]; packed[7] = (uint32_t)o1[3];
        break;
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
    case FMT_LUMINANCE:
        quantize_luminance(c0, c1, packed, q);
        best_format = FMT_LUMINANCE;
        break;
    case FMT_LUMINANCE_ALPHA:
        if (q <= Q160 && try_la_delta(c0, c1, packed, q)) {
            best_format = FMT_LUMINANCE_ALPHA_DELTA;
            break;
        }
        quantize_la(c0, c1, packed, q);
        best_format = FMT_LUMINANCE_ALPHA;
        break;
    default:
        best_format = FMT_LUMINANCE;
        quantize_luminance(c0, c1, packed, q);
        break;
    }
    return best_format;
}

/* LDR endpoint decode of pass15_unpack_color_endpoints.wgsl:307-391 (x257 applied by caller) */
void or_unpack_color_endpoints(int format, const uint32_t* v, int* o0, int* o1) {
    int in0[4], in1[4];
    switch (format) {
    case FMT_LUMINANCE: {
        int l0 = (int)v[0], l1 = (int)v[1];
        o0[0] = o0[1] = o0[2] = l0; o0[3] = 255;
        o1[0] = o1[1] = o1[2] = l1; o1[3] = 255;
        break;
    }
    case FMT_LUMINANCE_DELTA: {
        int v0 = (int)v[0], v1 = (int)v[1];
        int l0 = (v0 >> 2) | (v1 & 0xC0);
        int l1 = l0 + (v1 & 0x3F);
        if (l1 > 255) l1 = 255;
        o0[0] = o0[1] = o0[2] = l0; o0[3] = 255;
        o1[0] = o1[1] = o1[2] = l1; o1[3] = 255;
        break;
    }
    case FMT_LUMINANCE_ALPHA: {
        o0[0] = o0[1] = o0[2] = (int)v[0]; o0[3] = (int)v[2];
        o1[0] = o1[1] = o1[2] = (int)v[1]; o1[3] = (int)v[3];
        break;
    }
    case FMT_LUMINANCE_ALPHA_DELTA: {
        int l0 = (int)v[0], l1 = (int)v[1], a0 = (int)v[2], a1 = (int)v[3];
        l0 |= (l1 & 0x80) << 1;
        a0 |= (a1 & 0x80) << 1;
        l1 &= 0x7F; a1 &= 0x7F;
        if (l1 & 0x40) l1 -= 0x80;
        if (a1 & 0x40) a1 -= 0x80;
        l0 >>= 1; l1 >>= 1; a0 >>= 1; a1 >>= 1;
        l1 = l0 + l1; a1 = a0 + a1;
        l1 = clampi(l1, 0, 255); a1 = clampi(a1, 0, 255);
        o0[0] = o0[1] = o0[2] = l0; o0[3] = a0;
        o1[0] = o1[1] = o1[2] = l1; o1[3] = a1;
        break;
    }
    case FMT_RGB_SCALE: {
        int s = (int)v[3];
        int i0[4] = {(int)v[0], (int)v[1], (int)v[2], 255};
        memcpy(o1, i0, sizeof(i0));
        for (int c = 0; c < 4; c++) o0[c] = (i0[c] * s) >> 8;
        o0[3] = 255;
        break;
    }
    case FMT_RGB_SCALE_ALPHA: {
        int s = (int)v[3];
        int i0[4] = {(int)v[0], (int)v[1], (int)v[2], (int)v[4]};
        o1[0] = i0[0]; o1[1] = i0[1]; o1[2] = i0[2]; o1[3] = (int)v[5];
        for (int c = 0; c < 4; c++) o0[c] = (i0[c] * s) >> 8;
        o0[3] = (int)v[4];
        break;
    }
    case FMT_RGB:
    case FMT_RGB_DELTA:
        in0[0] = (int)v[0]; in0[1] = (int)v[2]; in0[2] = (int)v[4]; in0[3] = 0;
        in1[0] = (int)v[1]; in1[1] = (int)v[3]; in1[2] = (int)v[5]; in1[3] = 0;
        if (format == FMT_RGB) or_rgba_unpack(in0, in1, o0, o1);
        else or_rgba_delta_unpack(in0, in1, o0, o1);
        o0[3] = 255; o1[3] = 255;
        break;
    case FMT_RGBA:
    case FMT_RGBA_DELTA:
        in0[0] = (int)v[0]; in0[1] = (int)v[2]; in0[2] = (int)v[4]; in0[3] = (int)v[6];
        in1[0] = (int)v[1]; in1[1] = (int)v[3]; in1[2] = (int)v[5]; in1[3] = (int)v[7];
        if (format == FMT_RGBA) or_rgba_unpack(in0, in1, o0, o1);
        else or_rgba_delta_unpack(in0, in1, o0, o1);
        break;
    default: {
        int l0 = (int)v[0], l1 = (int)v[1];
        o0[0] = o0[1] = o0[2] = l0; o0[3] = 255;
        o1[0] = o1[1] = o1[2] = l1; o1[3] = 255;
        break;
    }
    }
}
