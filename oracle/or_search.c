/*
 * Oracle: per-block search (test infrastructure only).
 *
 * Restates ASTCEncoder::encode (code/compression_pipeline.cpp:323-535) and
 * the WGSL passes it dispatches, one partitioned block at a time, with every
 * workgroup reduction evaluated sequentially in index order.
 */
#include <stdlib.h>
#include <stdio.h>
#include "or_internal.h"
#include "or_color.h"

static const float CW[4] = {0.30f * 2.25f, 0.59f * 2.25f, 0.11f * 2.25f, 1.0f}; /* astc.h:60-63 */
static const int FREE_BITS[4] = {111, 111 - 4 - 10, 108 - 4 - 10, 105 - 4 - 10};  /* pass07:5 */
static const int STEPS_Q[12] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32};
#define REF_BATCH 920 /* astc.h:739 */

typedef struct {
    float px[OR_MAX_TEXELS][4];
    uint8_t tp[OR_MAX_TEXELS];
    int ppc[4];
    int partitioning_idx;
    int grayscale, constant_alpha;
} or_block;

typedef struct { float avg[4], dir[4], ep0[4], ep1[4]; } or_pep;

typedef struct {
    or_pep part[4];
    float weights[OR_MAX_TEXELS];
    float wes[OR_MAX_TEXELS];
    int is_const_wes;
    float min_weight_cutoff;
} or_ideal;

typedef struct {
    int bm_packed;       /* packed block mode index */
    int bm_trial;        /* valid-bm lookup index */
    float total_error;
    int quant_level, quant_level_mod;
    int formats_matched, final_quant_mode;
    int formats[4];
    uint32_t qw[OR_MAX_WEIGHTS];
    or_pep part[4];
    int final_formats[4];
    uint32_t packed[32];
} or_cand;

/* ---------- Cephes-style atan2f (single precision) ---------- */
static float atanf_poly(float x) {
    float sign = 1.0f;
    if (x < 0.0f) { sign = -1.0f; x = -x; }
    float y;
    if (x > 2.414213562373095f) { y = 1.5707963267948966f; x = -1.0f / x; }
    else if (x > 0.4142135623730950f) { y = 0.7853981633974483f; x = (x - 1.0f) / (x + 1.0f); }
    else y = 0.0f;
    float z = x * x;
    float p = 8.05374449538e-2f;
    p = p * z - 1.38776856032e-1f;
    p = p * z + 1.99777106478e-1f;
    p = p * z - 3.33329491539e-1f;
    y = y + (p * z * x + x);
    return sign * y;
}

float or_atan2f(float y, float x) {
    if (x != x || y != y) return x + y;
    if (x == 0.0f) {
        if (y > 0.0f) return 1.5707963267948966f;
        if (y < 0.0f) return -1.5707963267948966f;
        return 0.0f;
    }
    float w = 0.0f;
    if (x < 0.0f) w = (y < 0.0f) ? -3.141592653589793f : 3.141592653589793f;
    return w + atanf_poly(y / x);
}

/* ---------- pass01_ideal_endpoints_and_weights.wgsl:63-207 ---------- */
static void pass01(const or_desc* d, const or_block* b, or_ideal* o) {
    int tc = d->texel_count;
    float error_weight = (((CW[0] + CW[1]) + CW[2]) + CW[3]) / 4.0f;
    float sum[4][4] = {{0}}, count[4] = {0}, avg[4][4];
    for (int i = 0; i < tc; i++) {
        int p = b->tp[i];
        if (p < 4) {
            for (int c = 0; c < 4; c++) sum[p][c] += b->px[i][c];
            count[p] += 1.0f;
        }
    }
    for (int p = 0; p < 4; p++) {
        for (int c = 0; c < 4; c++) avg[p][c] = count[p] > 0.0f ? sum[p][c] / count[p] : 0.0f;
        memcpy(o->part[p].avg, avg[p], sizeof(avg[p]));
    }
    float dirv[4][4] = {{0}}, minp[4], maxp[4];
    for (int p = 0; p < 4; p++) { minp[p] = 1e10f; maxp[p] = -1e10f; }
    for (int i = 0; i < tc; i++) {
        int p = b->tp[i];
        if (p < 4)
            for (int c = 0; c < 4; c++) {
                float pd = wmax(b->px[i][c] - avg[p][c], 0.0f);
                dirv[p][c] += pd * pd;
            }
    }
    float dir[4][4];
    for (int p = 0; p < 4; p++) {
        normalize4(dirv[p], dir[p]);
        memcpy(o->part[p].dir, dir[p], sizeof(dir[p]));
    }
    for (int i = 0; i < tc; i++) {
        int p = b->tp[i];
        if (p < 4) {
            float df[4];
            for (int c = 0; c < 4; c++) df[c] = b->px[i][c] - avg[p][c];
            float proj = dot4(df, dir[p]);
            minp[p] = wmin(minp[p], proj);
            maxp[p] = wmax(maxp[p], proj);
        }
    }
    float p0len = 0.0f, lensq[4];
    int is_const = 1;
    for (int p = 0; p < 4; p++) {
        if (minp[p] >= maxp[p]) { minp[p] = 0.0f; maxp[p] = 1e-7f; }
        float len = maxp[p] - minp[p];
        float ls = len * len;
        lensq[p] = ls;
        int allzero = dir[p][0] == 0.0f && dir[p][1] == 0.0f && dir[p][2] == 0.0f && dir[p][3] == 0.0f;
        for (int c = 0; c < 4; c++) {
            if (allzero) { o->part[p].ep0[c] = avg[p][c]; o->part[p].ep1[c] = avg[p][c]; }
            else {
                o->part[p].ep0[c] = avg[p][c] + dir[p][c] * minp[p];
                o->part[p].ep1[c] = avg[p][c] + dir[p][c] * maxp[p];
            }
        }
        if (p == 0) p0len = ls;
        else is_const = is_const && (ls == p0len);
    }
    for (int i = 0; i < tc; i++) {
        int p = b->tp[i];
        if (p < 4) {
            float df[4];
            for (int c = 0; c < 4; c++) df[c] = b->px[i][c] - avg[p][c];
            float proj = dot4(df, dir[p]);
            float span = wmax(maxp[p] - minp[p], 1e-6f);
            o->weights[i] = wclamp((proj - minp[p]) / span, 0.0f, 1.0f);
            o->wes[i] = lensq[p] * error_weight;
        }
    }
    o->is_const_wes = is_const;
    float min_ep[4] = {10.0f, 10.0f, 10.0f, 10.0f};
    for (int p = 0; p < 4; p++)
        for (int c = 0; c < 4; c++) {
            float ep = (1.0f - o->part[p].ep0[c]) / ((o->part[p].ep1[c] - o->part[p].ep0[c]) + 1e-9f);
            if (ep > 0.5f && ep < min_ep[c]) min_ep[c] = ep;
        }
    o->min_weight_cutoff = wmin(min_ep[0], wmin(min_ep[1], wmin(min_ep[2], min_ep[3])));
}

static inline float wes_of(const or_ideal* ei, int t) { return ei->wes[ei->is_const_wes ? 0 : t]; }

/* ---------- pass02_decimated_weights.wgsl:132-270 ---------- */
static void pass02(const or_desc* d, const or_dm* di, const or_ideal* ei, float* dw) {
    int tc = di->texel_count, wc = di->weight_count;
    if (tc == wc) {
        for (int i = 0; i < wc; i++) dw[i] = ei->weights[i];
        return;
    }
    for (int w = 0; w < wc; w++) {
        int n = di->weight_texel_count[w];
        if (n == 0) { dw[w] = 0.0f; continue; }
        int off = di->w2t_off + di->weight_texels_offset[w];
        float ws = 0.0f, tcon = 1e-10f;
        for (int j = 0; j < n; j++) {
            int t = d->w2t_idx[off + j];
            float cw = d->w2t_contrib[off + j] * wes_of(ei, t);
            ws += ei->weights[t] * cw;
            tcon += cw;
        }
        dw[w] = ws / tcon;
    }
    float infill[OR_MAX_TEXELS];
    for (int i = 0; i < tc; i++) {
        int off = di->t2w_off + di->texel_weights_offset[i];
        float v = 0.0f;
        for (int j = 0; j < di->texel_weight_count[i]; j++) v += dw[d->t2w_idx[off + j]] * d->t2w_contrib[off + j];
        infill[i] = v;
    }
    float e0[OR_MAX_WEIGHTS], e1[OR_MAX_WEIGHTS];
    for (int w = 0; w < wc; w++) { e0[w] = 1e-10f; e1[w] = 0.0f; }
    for (int w = 0; w < wc; w++) {
        int n = di->weight_texel_count[w];
        int off = di->w2t_off + di->weight_texels_offset[w];
        for (int j = 0; j < n; j++) {
            int t = d->w2t_idx[off + j];
            float contrib = d->w2t_contrib[off + j];
            float scale = wes_of(ei, t) * contrib;
            e0[w] += contrib * scale;
            e1[w] += (infill[t] - ei->weights[t]) * scale;
        }
    }
    for (int w = 0; w < wc; w++) {
        float step = (e1[w] * -16.0f) / e0[w];
        dw[w] += wclamp(step, -0.25f, 0.25f);
    }
}

/* ---------- pass03_compute_angular_offsets.wgsl:65-116 ---------- */
static void pass03(const or_desc* d, const or_dm* di, const float* dw, float* ao) {
    int wc = di->weight_count;
    uint32_t is[OR_MAX_WEIGHTS];
    for (int i = 0; i < wc; i++) is[i] = wf2u(wround(wclamp(dw[i], 0.0f, 1.0f) * (1024.0f - 1.0f)));
    const float inv2pi = 1.0f / (2.0f * 3.14159265358979323846f);
    for (int s = 0; s < 16; s++) {
        float ax = 0.0f, ay = 0.0f;
        for (int j = 0; j < wc; j++) {
            ax += d->cos_table[is[j] * 16 + s];
            ay += d->sin_table[is[j] * 16 + s];
        }
        ao[s] = or_atan2f(ay, ax) * inv2pi;
    }
}

typedef struct { float lowest; int span; float error, cut_low, cut_high; } or_hl;

/* ---------- pass04_lowest_and_highest_weight.wgsl:116-227 ---------- */
static void pass04(const or_dm* di, const float* dw, const float* ao, or_hl* hl) {
    int wc = di->weight_count;
    float minw = 3.4e38f, maxw = -3.4e38f;
    for (int i = 0; i < wc; i++) {
        if (dw[i] < minw) minw = dw[i];
        if (dw[i] > maxw) maxw = dw[i];
    }
    for (int sp = 0; sp < di->max_angular_steps; sp++) {
        float rcp = (float)(sp + 1);
        float off = ao[sp];
        float minidx = wround(minw * rcp - off), maxidx = wround(maxw * rcp - off);
        float err = 0.0f, cl = 0.0f, ch = 0.0f;
        for (int j = 0; j < wc; j++) {
            float sval = dw[j] * rcp - off;
            float r = wround(sval);
            float diff = sval - r;
            err += diff * diff;
            cl += (r == minidx) ? 1.0f - (2.0f * diff) : 0.0f;
            ch += (r == maxidx) ? 1.0f + (2.0f * diff) : 0.0f;
        }
        int span = wf2i(maxidx - minidx + 1.0f);
        int hi = di->max_quant_steps + 3;
        span = span < 2 ? 2 : (span > hi ? hi : span);
        float ssize = 1.0f / rcp;
        float es = ssize * ssize;
        hl[sp].lowest = minidx;
        hl[sp].span = span;
        hl[sp].error = err * es;
        hl[sp].cut_low = cl * es;
        hl[sp].cut_high = ch * es;
    }
}

/* ---------- pass05_best_values_for_quant_levels.wgsl:84-175 ----------
 * low/high[q] for q <= max_quant_level; writes[q] tells which were produced. */
static void pass05(const or_dm* di, const or_hl* hl, const float* ao, float* low, float* high) {
    float rec[36][3];
    for (int i = 0; i < di->max_quant_steps + 4 && i < 36; i++) { rec[i][0] = 1e37f; rec[i][1] = -1.0f; rec[i][2] = 0.0f; }
    for (int i = 0; i < di->max_angular_steps; i++) {
        float fi = (float)i;
        int sp = hl[i].span;
        float e = hl[i].error, ecl = e + hl[i].cut_low, ech = e + hl[i].cut_high;
        float eclh = (e + hl[i].cut_low) + hl[i].cut_high;
        if (e < rec[sp][0]) { rec[sp][0] = e; rec[sp][1] = fi; rec[sp][2] = 0.0f; }
        if (ecl < rec[sp - 1][0]) { rec[sp - 1][0] = ecl; rec[sp - 1][1] = fi; rec[sp - 1][2] = 1.0f; }
        if (ech < rec[sp - 1][0]) { rec[sp - 1][0] = ech; rec[sp - 1][1] = fi; rec[sp - 1][2] = 0.0f; }
        if (eclh < rec[sp - 2][0]) { rec[sp - 2][0] = eclh; rec[sp - 2][1] = fi; rec[sp - 2][2] = 1.0f; }
    }
    for (int q = 0; q <= di->max_quant_level; q++) {
        int qs = STEPS_Q[q];
        int bsi = wf2i(rec[qs][1]);
        if (bsi < 0) bsi = 0;
        float lwi = hl[bsi].lowest + rec[qs][2];
        float hwi = (lwi + (float)qs) - 1.0f;
        float stepsize = 1.0f / (1.0f + (float)bsi);
        float off = ao[bsi];
        low[q] = (off + lwi) * stepsize;
        high[q] = (off + hwi) * stepsize;
    }
}

typedef struct {
    float error;
    int bitcount;
    uint8_t qw[OR_MAX_WEIGHTS];
} or_qres;

static const int Q2U_OFF[12] = {0, 2, 5, 9, 14, 20, 28, 38, 50, 66, 86, 110};

/* ---------- pass07_weights_and_error_for_bm.wgsl:183-332 ---------- */
static void pass07(const or_desc* d, const or_bm* bm, const or_dm* di, const or_ideal* ei, const float* dw,
                   float low, float high, int pc, or_qres* out) {
    (void)Q2U_OFF;
    memset(out->qw, 0, sizeof(out->qw));
    if (bm->quant_mode > 11) { out->error = 1e38f; out->bitcount = 0; return; }
    int bc = FREE_BITS[pc - 1] - bm->weight_bits;
    if (bc <= 0) { out->error = 1e38f; out->bitcount = 0; return; }
    if (high > 1.02f * ei->min_weight_cutoff) high = 1.0f;
    if (high <= low) { low = 0.0f; high = 1.0f; }
    int wc = di->weight_count, q = bm->quant_mode;
    float qlm1 = (float)(or_levels[q] - 1);
    float rscale = high - low;
    float scale = 1.0f / rscale;
    float slb = low * scale;
    rscale = rscale / 64.0f;
    float qwf[OR_MAX_WEIGHTS];
    uint8_t qwi[OR_MAX_WEIGHTS];
    for (int i = 0; i < wc; i++) {
        float ix = wclamp(dw[i] * scale - slb, 0.0f, 1.0f);
        float ix1 = ix * qlm1;
        uint32_t wl = wf2u(ix1);
        uint32_t wh = wl + 1u < wf2u(qlm1) ? wl + 1u : wf2u(qlm1);
        int il = or_quant_to_unquant[q][wl], ih = or_quant_to_unquant[q][wh];
        float fl = (float)il, fh = (float)ih;
        int use_high = (fl + fh) < (128.0f * ix);
        qwi[i] = (uint8_t)(use_high ? ih : il);
        qwf[i] = (use_high ? fh : fl) * rscale + low;
    }
    float err = 0.0f;
    for (int i = 0; i < di->texel_count; i++) {
        int off = di->t2w_off + di->texel_weights_offset[i];
        float v = 0.0f;
        for (int j = 0; j < di->texel_weight_count[i]; j++) v += qwf[d->t2w_idx[off + j]] * d->t2w_contrib[off + j];
        float diff = v - ei->weights[i];
        err += diff * diff * wes_of(ei, i);
    }
    out->error = err;
    out->bitcount = bc;
    /* only whole groups of four weights are stored (pass07:322-331) */
    for (int i = 0; i < (wc / 4) * 4; i++) out->qw[i] = qwi[i];
}

typedef struct {
    float rgb_scale_error, rgb_luma_error, luminance_error, alpha_drop_error;
    int can_offset, can_blue_contract;
} or_ece;

/* ---------- pass08_compute_encoding_choice_errors.wgsl:115-326 ---------- */
static void pass08(const or_desc* d, const or_block* b, const or_ideal* ei, int pc, or_ece* out) {
    float sxp[4][4] = {{0}}, syp[4][4] = {{0}}, szp[4][4] = {{0}};
    for (int i = 0; i < d->texel_count; i++) {
        int p = b->tp[i];
        if (p < pc) {
            float t[4];
            for (int c = 0; c < 4; c++) t[c] = b->px[i][c] - ei->part[p].avg[c];
            t[3] = 0.0f;
            if (t[0] > 0.0f) for (int c = 0; c < 4; c++) sxp[p][c] += t[c];
            if (t[1] > 0.0f) for (int c = 0; c < 4; c++) syp[p][c] += t[c];
            if (t[2] > 0.0f) for (int c = 0; c < 4; c++) szp[p][c] += t[c];
        }
    }
    float ub[4][4], ua[4][4], sb[4][4], lb[4], la[4][4], avg[4][4];
    float one[4] = {1.0f, 1.0f, 1.0f, 0.0f};
    normalize4(one, lb);
    for (int p = 0; p < pc; p++) {
        float* best = sxp[p];
        float bs = dot4(sxp[p], sxp[p]);
        float py = dot4(syp[p], syp[p]), pz = dot4(szp[p], szp[p]);
        if (py > bs) { best = syp[p]; bs = py; }
        if (pz > bs) { best = szp[p]; bs = pz; }
        for (int c = 0; c < 3; c++) avg[p][c] = ei->part[p].avg[c];
        avg[p][3] = 0.0f;
        normalize4(best, ub[p]);
        float s = dot4(avg[p], ub[p]);
        for (int c = 0; c < 4; c++) ua[p][c] = avg[p][c] - ub[p][c] * s;
        normalize4(avg[p], sb[p]);
        float sl = dot4(avg[p], lb);
        for (int c = 0; c < 4; c++) la[p][c] = avg[p][c] - lb[c] * sl;
    }
    float acc[4][5] = {{0}};
    for (int i = 0; i < d->texel_count; i++) {
        int p = b->tp[i];
        if (p < pc) {
            const float* rgb = b->px[i];
            float ad = b->px[i][3] - 65536.0f;
            acc[p][0] += ad * ad * CW[3];
            float dist[3], s;
            s = dot3(rgb, ub[p]);
            for (int c = 0; c < 3; c++) { dist[c] = (ua[p][c] + s * ub[p][c]) - rgb[c]; dist[c] = dist[c] * dist[c]; }
            acc[p][1] += dot3(dist, CW);
            s = dot3(rgb, sb[p]);
            for (int c = 0; c < 3; c++) { dist[c] = s * sb[p][c] - rgb[c]; dist[c] = dist[c] * dist[c]; }
            acc[p][2] += dot3(dist, CW);
            s = dot3(rgb, lb);
            for (int c = 0; c < 3; c++) { dist[c] = (la[p][c] + s * lb[c]) - rgb[c]; dist[c] = dist[c] * dist[c]; }
            acc[p][3] += dot3(dist, CW);
            for (int c = 0; c < 3; c++) { dist[c] = s * lb[c] - rgb[c]; dist[c] = dist[c] * dist[c]; }
            acc[p][4] += dot3(dist, CW);
        }
    }
    const float lim = 0.12f * 65536.0f;
    for (int p = 0; p < pc; p++) {
        out[p].rgb_scale_error = (acc[p][2] - acc[p][1]) * 0.7f;
        out[p].rgb_luma_error = (acc[p][3] - acc[p][1]) * 1.5f;
        out[p].luminance_error = (acc[p][4] - acc[p][1]) * 3.0f;
        out[p].alpha_drop_error = acc[p][0] * 3.0f;
        int co = 1;
        for (int c = 0; c < 3; c++) co = co && (fabsf(ei->part[p].ep1[c] - ei->part[p].ep0[c]) < lim);
        out[p].can_offset = co;
        out[p].can_blue_contract = !(b->grayscale == 1 && b->constant_alpha == 1);
    }
}

/* ---------- pass09_compute_color_error.wgsl:140-268 ---------- */
static void pass09(const or_block* b, const or_ideal* ei, const or_ece* ece, int pc, float err[4][21][4],
                   int fmt[4][21][4]) {
    static const int K[17] = {5, 7, 9, 11, 15, 19, 23, 31, 39, 47, 63, 79, 95, 127, 159, 191, 255};
    for (int p = 0; p < pc; p++) {
        const float* e0 = ei->part[p].ep0;
        const float* e1 = ei->part[p].ep1;
        float size = (float)b->ppc[p];
        float sre[4];
        for (int c = 0; c < 4; c++) {
            float h0 = wmax(e0[c] - 65535.0f, 0.0f), h1 = wmax(e1[c] - 65535.0f, 0.0f);
            float l0 = wmin(e0[c], 0.0f), l1 = wmin(e1[c], 0.0f);
            sre[c] = (((l0 * l0) + (l1 * l1)) + (h0 * h0)) + (h1 * h1);
        }
        float rgb_range = dot3(sre, CW) * 0.5f * size;
        float alpha_range = sre[3] * CW[3] * 0.5f * size;
        float base_rgb = ((CW[0] + CW[1]) + CW[2]) * size;
        float base_a = CW[3] * size;
        float bc_rgba = ece[p].can_blue_contract ? 0.625f : 1.0f;
        float oe_rgba0 = ece[p].can_offset ? 0.5f : 1.0f;
        float bc_rgb = ece[p].can_blue_contract ? 0.5f : 1.0f;
        float oe_rgb0 = ece[p].can_offset ? 0.25f : 1.0f;
        for (int q = 0; q < 4; q++) {
            for (int k = 0; k < 4; k++) err[p][q][k] = OR_ERR_DEFAULT;
            fmt[p][q][3] = FMT_RGBA; fmt[p][q][2] = FMT_RGB; fmt[p][q][1] = FMT_RGB_SCALE; fmt[p][q][0] = FMT_LUMINANCE;
        }
        for (int i = 0; i < 17; i++) {
            int q = i + 4;
            float oe_rgba = q >= 19 ? 1.0f : oe_rgba0, oe_rgb = q >= 19 ? 1.0f : oe_rgb0;
            float bqe = (65536.0f * 65536.0f / 18.0f) / (float)(K[i] * K[i]);
            float qe_rgb = base_rgb * bqe;
            float qe_rgba = (base_rgb + base_a) * bqe;
            err[p][q][3] = ((qe_rgba * bc_rgba * oe_rgba) + rgb_range) + alpha_range;
            fmt[p][q][3] = FMT_RGBA;
            float full = ((qe_rgb * bc_rgb * oe_rgb) + rgb_range) + ece[p].alpha_drop_error;
            float rgbs_a = ((qe_rgba + ece[p].rgb_scale_error) + rgb_range) + alpha_range;
            if (rgbs_a < full) { err[p][q][2] = rgbs_a; fmt[p][q][2] = FMT_RGB_SCALE_ALPHA; }
            else { err[p][q][2] = full; fmt[p][q][2] = FMT_RGB; }
            float rgbs = ((qe_rgb + rgb_range) + ece[p].alpha_drop_error) + ece[p].rgb_scale_error;
            float lumal = ((qe_rgba + rgb_range) + alpha_range) + ece[p].luminance_error;
            if (rgbs < lumal) { err[p][q][1] = rgbs; fmt[p][q][1] = FMT_RGB_SCALE; }
            else { err[p][q][1] = lumal; fmt[p][q][1] = FMT_LUMINANCE_ALPHA; }
            err[p][q][0] = ((qe_rgb + rgb_range) + ece[p].alpha_drop_error) + ece[p].luminance_error;
            fmt[p][q][0] = FMT_LUMINANCE;
        }
    }
}

typedef struct { float error; int formats[4]; } or_comb;

/* ---------- pass10_color_combinations_for_quant_{2,3,4}part.wgsl ---------- */
static void pass10(int pc, float err[4][21][4], int fmt[4][21][4], or_comb comb[21][13]) {
    for (int q = 0; q < 21; q++)
        for (int k = 0; k < 13; k++) { comb[q][k].error = OR_ERR_DEFAULT; memset(comb[q][k].formats, 0, sizeof(comb[q][k].formats)); }
    for (int q = 4; q < 21; q++) {
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                int lo2 = i < j ? i : j, hi2 = i > j ? i : j;
                if (hi2 - lo2 > 1) continue;
                if (pc == 2) {
                    float te = wmin(err[0][q][i] + err[1][q][j], 1e10f);
                    or_comb* o = &comb[q][i + j];
                    if (te < o->error) { o->error = te; o->formats[0] = fmt[0][q][i]; o->formats[1] = fmt[1][q][j]; o->formats[2] = 0; o->formats[3] = 0; }
                    continue;
                }
                for (int k = 0; k < 4; k++) {
                    int lo3 = k < lo2 ? k : lo2, hi3 = k > hi2 ? k : hi2;
                    if (hi3 - lo3 > 1) continue;
                    if (pc == 3) {
                        float te = wmin((err[0][q][i] + err[1][q][j]) + err[2][q][k], 1e10f);
                        or_comb* o = &comb[q][i + j + k];
                        if (te < o->error) { o->error = te; o->formats[0] = fmt[0][q][i]; o->formats[1] = fmt[1][q][j]; o->formats[2] = fmt[2][q][k]; o->formats[3] = 0; }
                        continue;
                    }
                    for (int l = 0; l < 4; l++) {
                        int lo4 = l < lo3 ? l : lo3, hi4 = l > hi3 ? l : hi3;
                        if (hi4 - lo4 > 1) continue;
                        float te = wmin(((err[0][q][i] + err[1][q][j]) + err[2][q][k]) + err[3][q][l], 1e10f);
                        or_comb* o = &comb[q][i + j + k + l];
                        if (te < o->error) { o->error = te; o->formats[0] = fmt[0][q][i]; o->formats[1] = fmt[1][q][j]; o->formats[2] = fmt[2][q][k]; o->formats[3] = fmt[3][q][l]; }
                    }
                }
            }
    }
}

typedef struct { float total_error; int ql, ql_mod; int formats[4]; } or_ccr;

/* ---------- pass11_best_color_combination_for_mode_{1,2,3,4}part.wgsl ---------- */
static void pass11(int pc, const or_qres* qr, float err[4][21][4], int fmt[4][21][4], or_comb comb[21][13],
                   or_ccr* out) {
    float we = qr->error;
    int bits = qr->bitcount;
    if (we >= OR_ERR_DEFAULT) { out->total_error = OR_ERR_DEFAULT; return; }
    memset(out->formats, 0, sizeof(out->formats));
    if (pc == 1) {
        int best_idx = 0;
        float best = OR_ERR_DEFAULT;
        for (int ic = 1; ic <= 4; ic++) {
            int ql = or_quant_mode_table[ic][bits];
            if (ql < 4) continue;
            float e = err[0][ql][ic - 1];
            if (e < best) { best = e; best_idx = ic; }
        }
        int fql = or_quant_mode_table[best_idx][bits];
        int f = FMT_LUMINANCE;
        if (fql >= 4) f = fmt[0][fql][best_idx - 1];
        out->total_error = best + we;
        out->ql = fql; out->ql_mod = fql;
        out->formats[0] = f;
        return;
    }
    int start = pc, end = pc == 2 ? 8 : 9, modbits = pc == 2 ? 2 : (pc == 3 ? 5 : 8);
    int best_idx = 0;
    float best = OR_ERR_DEFAULT;
    for (int ic = start; ic <= end; ic++) {
        int ql = or_quant_mode_table[ic][bits];
        if (ql < 4) break;
        float e = comb[ql][ic - start].error;
        if (e < best) { best = e; best_idx = ic; }
    }
    int fql = or_quant_mode_table[best_idx][bits];
    int fqlm = or_quant_mode_table[best_idx][bits + modbits];
    if (fql >= 4) memcpy(out->formats, comb[fql][best_idx - start].formats, sizeof(out->formats));
    out->total_error = best + we;
    out->ql = fql; out->ql_mod = fqlm;
}

/* ---------- pass13_recompute_ideal_endpoints.wgsl:182-393 ---------- */
static void pass13(const or_desc* d, const or_block* b, int pc, or_cand* c, float rgbs[4][4]) {
    const or_bm* bm = &d->bm[c->bm_packed];
    const or_dm* di = &d->dm[bm->decimation_mode];
    float decw[OR_MAX_WEIGHTS], undec[OR_MAX_TEXELS];
    for (int i = 0; i < di->weight_count; i++) decw[i] = (float)c->qw[i] / 64.0f;
    for (int i = 0; i < di->texel_count; i++) {
        if (di->texel_count == di->weight_count) undec[i] = decw[i];
        else {
            int off = di->t2w_off + di->texel_weights_offset[i];
            float v = 0.0f;
            for (int j = 0; j < di->texel_weight_count[i]; j++) v += decw[d->t2w_idx[off + j]] * d->t2w_contrib[off + j];
            undec[i] = v;
        }
    }
    float wmin1[4], wmax1[4], ls[4], ms[4], rs[4], smin[4], smax[4], cx[4][4], cy[4][4], sv[4][2], sdir[4][3];
    for (int p = 0; p < pc; p++) {
        wmin1[p] = 1.0f; wmax1[p] = 0.0f; ls[p] = ms[p] = rs[p] = 0.0f; smin[p] = 1e10f; smax[p] = 0.0f;
        for (int k = 0; k < 4; k++) { cx[p][k] = 0.0f; cy[p][k] = 0.0f; }
        sv[p][0] = sv[p][1] = 0.0f;
        float size = (float)b->ppc[p];
        float rsum[3], rws[3];
        for (int k = 0; k < 3; k++) {
            rsum[k] = (c->part[p].avg[k] * size) * CW[k];
            rws[k] = wmax(CW[k] * size, 1e-17f);
        }
        float q3[3] = {rsum[0] / rws[0], rsum[1] / rws[1], rsum[2] / rws[2]};
        normalize3(q3, sdir[p]);
    }
    const float lsw = (CW[0] + CW[1]) + CW[2];
    for (int i = 0; i < di->texel_count; i++) {
        int p = b->tp[i];
        const float* rgba = b->px[i];
        float w = undec[i];
        if (w < wmin1[p]) wmin1[p] = w;
        if (w > wmax1[p]) wmax1[p] = w;
        float sc = dot3(rgba, sdir[p]);
        if (sc < smin[p]) smin[p] = sc;
        if (sc > smax[p]) smax[p] = sc;
        float om = 1.0f - w;
        ls[p] += om * om;
        ms[p] += om * w;
        rs[p] += w * w;
        for (int k = 0; k < 4; k++) {
            float cyk = rgba[k] * w;
            cy[p][k] += cyk;
            cx[p][k] += rgba[k] - cyk;
        }
        sv[p][0] += om * sc * lsw;
        sv[p][1] += w * sc * lsw;
    }
    for (int p = 0; p < pc; p++) {
        float colx[4], coly[4], lv[4], mv[4], rv[4];
        for (int k = 0; k < 4; k++) {
            colx[k] = cx[p][k] * CW[k]; coly[k] = cy[p][k] * CW[k];
            lv[k] = ls[p] * CW[k]; mv[k] = ms[p] * CW[k]; rv[k] = rs[p] * CW[k];
        }
        float lx = ls[p] * lsw, ly = ms[p] * lsw, lz = rs[p] * lsw;
        float scalediv = wclamp(smin[p] / wmax(smax[p], 1e-10f), 0.0f, 1.0f);
        float sds[3] = {sdir[p][0] * smax[p], sdir[p][1] * smax[p], sdir[p][2] * smax[p]};
        rgbs[p][0] = sds[0]; rgbs[p][1] = sds[1]; rgbs[p][2] = sds[2]; rgbs[p][3] = scalediv;
        or_pep* pp = &c->part[p];
        if (wmin1[p] >= wmax1[p] * 0.999f) {
            float size = (float)b->ppc[p];
            for (int k = 0; k < 4; k++) {
                float rws = wmax(CW[k] * size, 1e-17f);
                float ac = (colx[k] + coly[k]) / rws;
                if (ac == ac) { pp->ep0[k] = ac; pp->ep1[k] = ac; }
            }
            rgbs[p][3] = 1.0f;
        } else {
            for (int k = 0; k < 4; k++) {
                float det = (lv[k] * rv[k]) - (mv[k] * mv[k]);
                float rdet = 1.0f / det;
                float mss = ((lv[k] * lv[k]) + 2.0f * (mv[k] * mv[k])) + (rv[k] * rv[k]);
                float e0 = (rv[k] * colx[k] - mv[k] * coly[k]) * rdet;
                float e1 = (lv[k] * coly[k] - mv[k] * colx[k]) * rdet;
                if (fabsf(det) > (mss * 1e-4f) && e0 == e0 && e1 == e1) { pp->ep0[k] = e0; pp->ep1[k] = e1; }
            }
            float lsdet = (lx * lz) - (ly * ly);
            float lsrdet = 1.0f / lsdet;
            float lsmss = ((lx * lx) + 2.0f * (ly * ly)) + (lz * lz);
            float se0 = (lz * sv[p][0] - ly * sv[p][1]) * lsrdet;
            float se1 = (lx * sv[p][1] - ly * sv[p][0]) * lsrdet;
            if ((fabsf(lsdet) > (lsmss * 1e-4f)) && se0 == se0 && se1 == se1 && se0 < se1) {
                rgbs[p][0] = sdir[p][0] * se1; rgbs[p][1] = sdir[p][1] * se1; rgbs[p][2] = sdir[p][2] * se1;
                rgbs[p][3] = se0 / se1;
            }
        }
    }
}

/* ---------- pass14_pack_color_endpoints.wgsl:1226-1304 ---------- */
static void pass14(int pc, or_cand* c, float rgbs[4][4]) {
    int all_same = c->quant_level != c->quant_level_mod;
    int fn[4] = {0, 0, 0, 0};
    uint32_t vn[32];
    memset(vn, 0, sizeof(vn));
    for (int p = 0; p < pc; p++) {
        uint32_t v[8];
        int f = or_pack_color_endpoints(c->part[p].ep0, c->part[p].ep1, rgbs[p], c->formats[p], c->quant_level, v);
        fn[p] = f;
        for (int i = 0; i < 8; i++) vn[p * 8 + i] = v[i];
        if (p > 0) all_same = all_same && (f == fn[0]);
    }
    int matched = 0;
    if (pc >= 2 && all_same) {
        int fm[4] = {0, 0, 0, 0}, same_mod = 1;
        uint32_t vm[32];
        memset(vm, 0, sizeof(vm));
        for (int p = 0; p < pc; p++) {
            uint32_t v[8];
            int f = or_pack_color_endpoints(c->part[p].ep0, c->part[p].ep1, rgbs[p], c->formats[p], c->quant_level_mod, v);
            fm[p] = f;
            for (int i = 0; i < 8; i++) vm[p * 8 + i] = v[i];
            if (p > 0 && f != fm[0]) { same_mod = 0; break; }
        }
        if (same_mod) { matched = 1; memcpy(fn, fm, sizeof(fn)); memcpy(vn, vm, sizeof(vn)); }
    }
    c->formats_matched = matched;
    c->final_quant_mode = matched ? c->quant_level_mod : c->quant_level;
    memcpy(c->final_formats, fn, sizeof(fn));
    memcpy(c->packed, vn, sizeof(vn));
}

/* ---------- pass15_unpack_color_endpoints.wgsl:307-391 ---------- */
static void pass15(int pc, const or_cand* c, int ue0[4][4], int ue1[4][4]) {
    for (int p = 0; p < pc; p++) {
        int o0[4], o1[4];
        or_unpack_color_endpoints(c->final_formats[p], c->packed + p * 8, o0, o1);
        for (int k = 0; k < 4; k++) { ue0[p][k] = o0[k] * 257; ue1[p][k] = o1[k] * 257; }
    }
}

/* ---------- pass16_realign_weights.wgsl:211-367 ---------- */
static void pass16(const or_desc* d, const or_block* b, int pc, or_cand* c, int ue0[4][4], int ue1[4][4]) {
    const or_bm* bm = &d->bm[c->bm_packed];
    const or_dm* di = &d->dm[bm->decimation_mode];
    float base[4][4], offs[4][4];
    for (int p = 0; p < pc; p++)
        for (int k = 0; k < 4; k++) {
            base[p][k] = (float)ue0[p][k];
            offs[p][k] = (float)(ue1[p][k] - ue0[p][k]) * (1.0f / 64.0f);
        }
    float uq[OR_MAX_WEIGHTS];
    for (int i = 0; i < di->weight_count; i++) uq[i] = (float)c->qw[i];
    int q = bm->quant_mode;
    if (di->weight_count == di->texel_count) {
        for (int t = 0; t < di->texel_count; t++) {
            int uqw = (int)c->qw[t];
            uint32_t pn = or_prev_next[q][uqw];
            int down = (int)(pn & 0xFF), up = (int)((pn >> 8) & 0xFF);
            float wb = (float)uqw, dd = (float)(down - uqw), ud = (float)(up - uqw);
            int p = b->tp[t];
            float diff[4], dfd[4], dfu[4];
            for (int k = 0; k < 4; k++) {
                float col = base[p][k] + offs[p][k] * wb;
                diff[k] = col - b->px[t][k];
                dfd[k] = diff[k] + offs[p][k] * dd;
                dfu[k] = diff[k] + offs[p][k] * ud;
            }
            float s0[4], s1[4], s2[4];
            for (int k = 0; k < 4; k++) { s0[k] = diff[k] * diff[k]; s1[k] = dfd[k] * dfd[k]; s2[k] = dfu[k] * dfu[k]; }
            float eb = dot4(s0, CW), ed = dot4(s1, CW), eu = dot4(s2, CW);
            if ((eu < eb) && (eu < ed) && (uqw < 64)) c->qw[t] = (uint32_t)up;
            else if ((ed < eb) && (uqw > 0)) c->qw[t] = (uint32_t)down;
        }
        return;
    }
    for (int we = 0; we < di->weight_count; we++) {
        int uqw = (int)c->qw[we];
        uint32_t pn = or_prev_next[q][uqw];
        float down = (float)(pn & 0xFF), up = (float)((pn >> 8) & 0xFF), wbase = (float)uqw;
        float ddn = down - wbase, dup = up - wbase;
        float ebv[4] = {0}, edv[4] = {0}, euv[4] = {0};
        int n = di->weight_texel_count[we];
        int woff = di->w2t_off + di->weight_texels_offset[we];
        for (int te = 0; te < n; te++) {
            int t = d->w2t_idx[woff + te];
            int toff = di->t2w_off + di->texel_weights_offset[t];
            float wbt = 0.0f;
            for (int j = 0; j < di->texel_weight_count[t]; j++) wbt += uq[d->t2w_idx[toff + j]] * d->t2w_contrib[toff + j];
            float twb = 0.0f;
            for (int k = 0; k < di->texel_weight_count[t]; k++)
                if (d->t2w_idx[toff + k] == we) { twb = d->t2w_contrib[toff + k]; break; }
            float wdd = ddn * twb, wud = dup * twb;
            int p = b->tp[t];
            for (int k = 0; k < 4; k++) {
                float col = base[p][k] + offs[p][k] * wbt;
                float diff = col - b->px[t][k];
                float dd = diff + offs[p][k] * wdd;
                float du = diff + offs[p][k] * wud;
                ebv[k] += diff * diff;
                edv[k] += dd * dd;
                euv[k] += du * du;
            }
        }
        float eb = dot4(ebv, CW), ed = dot4(edv, CW), eu = dot4(euv, CW);
        if ((eu < eb) && (eu < ed) && (uqw < 64)) { c->qw[we] = (uint32_t)up; uq[we] = up; }
        else if ((ed < eb) && (uqw > 0)) { c->qw[we] = (uint32_t)down; uq[we] = down; }
    }
}

/* ---------- pass17_compute_final_error.wgsl:149-231 ---------- */
static float pass17(const or_desc* d, const or_block* b, int pc, const or_cand* c, int ue0[4][4], int ue1[4][4]) {
    const or_bm* bm = &d->bm[c->bm_packed];
    const or_dm* di = &d->dm[bm->decimation_mode];
    float decw[OR_MAX_WEIGHTS];
    for (int i = 0; i < di->weight_count; i++) decw[i] = (float)c->qw[i] / 64.0f;
    float total = 0.0f;
    for (int i = 0; i < di->texel_count; i++) {
        float w;
        if (di->texel_count == di->weight_count) w = decw[i];
        else {
            int off = di->t2w_off + di->texel_weights_offset[i];
            w = 0.0f;
            for (int j = 0; j < di->texel_weight_count[i]; j++) w += decw[d->t2w_idx[off + j]] * d->t2w_contrib[off + j];
        }
        int p = b->tp[i];
        if (p < pc) {
            int w1 = wf2i(wround(w * 64.0f)), w0 = 64 - w1;
            float dd[4];
            for (int k = 0; k < 4; k++) {
                int col = (ue0[p][k] * w0 + ue1[p][k] * w1 + 32) >> 6;
                float df = fabsf(b->px[i][k] - (float)col);
                df = wmin(df, 1e15f);
                dd[k] = df * df;
            }
            float e = dot4(dd, CW);
            total += wmin(e, OR_ERR_DEFAULT);
        }
    }
    return total;
}

/* ---------- passes 01-18 for one partitioned block ---------- */
static void search_pb(const or_desc* d, const or_block* b, int pc, int pb_in_batch, or_symbolic* out, or_stats* st) {
    or_ideal ei;
    pass01(d, b, &ei);
    int ndm = d->ndm, nbm = d->nbm;
    float* dws = (float*)malloc(sizeof(float) * (size_t)ndm * OR_MAX_WEIGHTS);
    float* lows = (float*)calloc((size_t)ndm * 13, sizeof(float));
    float* highs = (float*)calloc((size_t)ndm * 13, sizeof(float));
    /* pass05 output is allocated as 8 floats per decimation-mode trial for
     * 920 * 4 partitioned blocks but indexed with stride 13
     * (shaders_init.cpp:1176 vs pass05:171); accesses past the end read 0. */
    const int64_t alloc = (int64_t)8 * REF_BATCH * 4 * ndm;
    for (int m = 0; m < ndm; m++) {
        const or_dm* di = &d->dm[d->valid_dm[m]];
        float* dw = dws + (size_t)m * OR_MAX_WEIGHTS;
        float ao[16];
        or_hl hl[16];
        float lo[13], hi[13];
        pass02(d, di, &ei, dw);
        pass03(d, di, dw, ao);
        pass04(di, dw, ao, hl);
        pass05(di, hl, ao, lo, hi);
        int64_t t = (int64_t)pb_in_batch * ndm + m;
        for (int q = 0; q <= di->max_quant_level; q++)
            if (13 * t + q < alloc) { lows[m * 13 + q] = lo[q]; highs[m * 13 + q] = hi[q]; }
    }
    or_ece ece[4];
    pass08(d, b, &ei, pc, ece);
    static __thread float cerr[4][21][4];
    static __thread int cfmt[4][21][4];
    static __thread or_comb comb[21][13];
    pass09(b, &ei, ece, pc, cerr, cfmt);
    if (pc > 1) pass10(pc, cerr, cfmt, comb);
    or_qres* qr = (or_qres*)malloc(sizeof(or_qres) * (size_t)nbm);
    or_ccr* ccr = (or_ccr*)malloc(sizeof(or_ccr) * (size_t)nbm);
    for (int k = 0; k < nbm; k++) {
        const or_bm* bm = &d->bm[d->valid_bm[k]];
        int m = d->valid_bm_dml[k];
        /* pass06_remap_low_and_high_values.wgsl:72-98: quant <= 12 always here */
        float lo = lows[m * 13 + bm->quant_mode], hi = highs[m * 13 + bm->quant_mode];
        pass07(d, bm, &d->dm[bm->decimation_mode], &ei, dws + (size_t)m * OR_MAX_WEIGHTS, lo, hi, pc, &qr[k]);
        pass11(pc, &qr[k], cerr, cfmt, comb, &ccr[k]);
    }
    /* pass12_find_top_N_candidates.wgsl:121-191 */
    float te[8];
    int ti[8];
    for (int i = 0; i < 8; i++) { te[i] = OR_ERR_DEFAULT; ti[i] = 0; }
    for (int k = 0; k < nbm; k++) {
        float e = ccr[k].total_error;
        if (e < te[7]) {
            te[7] = e; ti[7] = k;
            for (int j = 7; j > 0; j--)
                if (te[j] < te[j - 1]) {
                    float fe = te[j]; te[j] = te[j - 1]; te[j - 1] = fe;
                    int fi = ti[j]; ti[j] = ti[j - 1]; ti[j - 1] = fi;
                }
        }
    }
    or_cand best_top[8];
    float best_err[8];
    int ncand = 0;
    for (int w = 0; w < 8; w++) {
        best_err[w] = OR_ERR_DEFAULT;
        if (!(te[w] < OR_ERR_DEFAULT)) continue;
        ncand++;
        or_cand c;
        memset(&c, 0, sizeof(c));
        int k = ti[w];
        c.bm_packed = d->valid_bm[k];
        c.bm_trial = k;
        c.total_error = ccr[k].total_error;
        c.quant_level = ccr[k].ql;
        c.quant_level_mod = ccr[k].ql_mod;
        memcpy(c.formats, ccr[k].formats, sizeof(c.formats));
        memcpy(c.part, ei.part, sizeof(c.part));
        for (int i = 0; i < OR_MAX_WEIGHTS; i++) c.qw[i] = qr[k].qw[i];
        /* refinement loop, compression_pipeline.cpp:473-484 */
        float rgbs[4][4];
        int ue0[4][4], ue1[4][4];
        for (int a = 0; a < 6; a++) {
            pass13(d, b, pc, &c, rgbs);
            pass14(pc, &c, rgbs);
            pass15(pc, &c, ue0, ue1);
            if (a == 0) {
                float e = pass17(d, b, pc, &c, ue0, ue1);
                c.total_error = e;
                if (e < best_err[w]) { best_err[w] = e; best_top[w] = c; }
            }
            pass16(d, b, pc, &c, ue0, ue1);
            float e = pass17(d, b, pc, &c, ue0, ue1);
            c.total_error = e;
            if (e < best_err[w]) { best_err[w] = e; best_top[w] = c; }
        }
    }
    if (st && ncand < 8) st->short_candidate_lists++;
    /* pass18_pick_best_candidate.wgsl:109-141 */
    float be = OR_ERR_DEFAULT;
    int bi = -1;
    for (int w = 0; w < 8; w++)
        if (best_err[w] < be) { be = best_err[w]; bi = w; }
    memset(out, 0, sizeof(*out));
    out->errorval = be;
    out->partition_count = (uint32_t)pc;
    out->partition_index = (uint32_t)b->partitioning_idx;
    if (bi >= 0) {
        const or_cand* c = &best_top[bi];
        out->block_mode_index = (uint32_t)d->bm[c->bm_packed].mode_index;
        out->quant_mode = (uint32_t)c->final_quant_mode;
        out->partition_formats_matched = (uint32_t)c->formats_matched;
        for (int i = 0; i < 4; i++) out->partition_formats[i] = (uint32_t)c->final_formats[i];
        for (int i = 0; i < 32; i++) out->packed_color_values[i] = c->packed[i];
        for (int i = 0; i < 64; i++) out->quantized_weights[i] = c->qw[i];
    }
    free(dws); free(lows); free(highs); free(qr); free(ccr);
}

/* ---------- SplitImageIntoBlocks, compression_pipeline.cpp:7-86 ---------- */
static void load_block(const or_desc* d, const uint8_t* rgba, int w, int h, int bx, int by, or_block* b) {
    float amin = 1e38f, amax = 0.0f;
    int gray = 1, i = 0;
    for (int dy = 0; dy < d->ydim; dy++) {
        int y = by * d->ydim + dy;
        if (y > h - 1) y = h - 1;
        for (int dx = 0; dx < d->xdim; dx++) {
            int x = bx * d->xdim + dx;
            if (x > w - 1) x = w - 1;
            const uint8_t* p = rgba + ((size_t)y * w + x) * 4;
            for (int c = 0; c < 4; c++) b->px[i][c] = ((float)p[c] / 255.0f) * 65536.0f;
            amax = fmaxf(amax, b->px[i][3]);
            amin = fminf(amin, b->px[i][3]);
            gray = gray && (b->px[i][0] == b->px[i][1] && b->px[i][0] == b->px[i][2]);
            b->tp[i] = 0;
            i++;
        }
    }
    b->grayscale = gray;
    b->constant_alpha = amax == amin;
    b->ppc[0] = d->texel_count; b->ppc[1] = b->ppc[2] = b->ppc[3] = 0;
    b->partitioning_idx = 0;
}

/* ---------- compat partition candidate stage: passes 001-007 ---------- */

static float dist_sq(const float* a, const float* b) {
    float d2[4];
    for (int c = 0; c < 4; c++) { float t = a[c] - b[c]; d2[c] = t * t; }
    return dot4(d2, CW);
}

/* pass001-003: texel assignments after 4 k-means iterations */
static void kmeans(const or_desc* d, const or_block* b, int pc, uint8_t* assign) {
    int tc = d->texel_count;
    float centers[4][4], dist[OR_MAX_TEXELS];
    memset(centers, 0, sizeof(centers));
    memcpy(centers[0], b->px[145897 % tc], sizeof(centers[0]));
    for (int t = 0; t < tc; t++) dist[t] = dist_sq(centers[0], b->px[t]);
    for (int p = 1; p < pc; p++) {
        uint32_t mx = 0;
        for (int t = 0; t < tc; t++) { uint32_t u; memcpy(&u, &dist[t], 4); if (u > mx) mx = u; }
        uint32_t idx = 0;
        for (int t = 0; t < tc; t++) { uint32_t u; memcpy(&u, &dist[t], 4); if (u == mx) idx = (uint32_t)t; }
        memcpy(centers[p], b->px[idx], sizeof(centers[p]));
        if (p < pc - 1)
            for (int t = 0; t < tc; t++) dist[t] = wmin(dist[t], dist_sq(centers[p], b->px[t]));
    }
    for (int it = 0; it < 4; it++) {
        int counts[4] = {0, 0, 0, 0};
        for (int t = 0; t < tc; t++) {
            float bd = 1e30f;
            int bp = 0;
            for (int p = 0; p < pc; p++) {
                float dd = dist_sq(b->px[t], centers[p]);
                if (dd < bd) { bd = dd; bp = p; }
            }
            assign[t] = (uint8_t)bp;
            counts[bp]++;
        }
        int problem = 0;
        for (int p = 0; p < pc; p++) if (counts[p] == 0) problem = 1;
        if (problem) for (int p = 0; p < pc; p++) assign[p] = (uint8_t)p;
        uint32_t sums[4][4] = {{0}}, cnt[4] = {0};
        for (int t = 0; t < tc; t++) {
            int p = assign[t];
            for (int c = 0; c < 4; c++) sums[p][c] += wf2u(b->px[t][c]);
            cnt[p]++;
        }
        for (int p = 0; p < pc; p++) {
            for (int c = 0; c < 4; c++) centers[p][c] = cnt[p] > 0 ? (float)sums[p][c] / (float)cnt[p] : 0.0f;
        }
    }
}

static int popc64(uint64_t v) { return __builtin_popcountll(v); }

/* pass004_count_partition_mismatch.wgsl:71-186 for one partitioning */
static uint32_t mismatch(const or_desc* d, const uint8_t* assign, int pc, int pidx) {
    uint64_t a[4] = {0, 0, 0, 0};
    for (int l = 0; l < 64; l++) a[assign[d->kmeans_texels[l]]] |= 1ull << l;
    const uint64_t* bb = d->coverage[pc - 2][pidx];
#define P(i, j) popc64(a[i] ^ bb[j])
#define MIN(x, y) ((x) < (y) ? (x) : (y))
    if (pc == 2) {
        int v1 = P(0, 0) + P(1, 1), v2 = P(0, 1) + P(1, 0);
        return (uint32_t)(MIN(v1, v2) / 2);
    }
    if (pc == 3) {
        int v0 = MIN(P(1, 1) + P(2, 2), P(1, 2) + P(2, 1)) + P(0, 0);
        int v1 = MIN(P(1, 0) + P(2, 2), P(1, 2) + P(2, 0)) + P(0, 1);
        int v2 = MIN(P(1, 0) + P(2, 1), P(1, 1) + P(2, 0)) + P(0, 2);
        return (uint32_t)(MIN(v0, MIN(v1, v2)) / 2);
    }
    int mx23 = MIN(P(2, 2) + P(3, 3), P(2, 3) + P(3, 2));
    int mx13 = MIN(P(2, 1) + P(3, 3), P(2, 3) + P(3, 1));
    int mx12 = MIN(P(2, 1) + P(3, 2), P(2, 2) + P(3, 1));
    int mx03 = MIN(P(2, 0) + P(3, 3), P(2, 3) + P(3, 0));
    int mx02 = MIN(P(2, 0) + P(3, 2), P(2, 2) + P(3, 0));
    int mx01 = MIN(P(2, 1) + P(3, 0), P(2, 0) + P(3, 1));
    int v0 = P(0, 0) + MIN(P(1, 1) + mx23, MIN(P(1, 2) + mx13, P(1, 3) + mx12));
    int v1 = P(0, 1) + MIN(P(1, 0) + mx23, MIN(P(1, 2) + mx03, P(1, 3) + mx02));
    int v2 = P(0, 2) + MIN(P(1, 1) + mx03, MIN(P(1, 0) + mx13, P(1, 3) + mx01));
    int v3 = P(0, 3) + MIN(P(1, 1) + mx02, MIN(P(1, 2) + mx01, P(1, 0) + mx12));
    return (uint32_t)(MIN(v0, MIN(v1, MIN(v2, v3))) / 2);
#undef P
#undef MIN
}

/* pass005_partition_ordering.wgsl:45-94: the partial bitonic network */
static void ordering(int nsel, uint32_t score0, uint32_t* ord128) {
    static __thread uint32_t sc[1024], ix[1024];
    for (int i = 0; i < 1024; i++) { ix[i] = (uint32_t)i; sc[i] = i < nsel ? (i == 0 ? score0 : 0u) : 0xFFFFFFFFu; }
    for (int k = 2; k <= 1024; k *= 2)
        for (int j = k / 2; j > 0; j /= 2)
            for (int l = 0; l < 256; l++) {
                int pt = l ^ j;
                if (pt > l) {
                    int asc = (l & k) == 0;
                    int sw = (sc[l] > sc[pt]) || (sc[l] == sc[pt] && ix[l] > ix[pt]);
                    if (sw == asc) {
                        uint32_t t = sc[l]; sc[l] = sc[pt]; sc[pt] = t;
                        t = ix[l]; ix[l] = ix[pt]; ix[pt] = t;
                    }
                }
            }
    for (int i = 0; i < 128; i++) ord128[i] = ix[i];
}

/* pass006_evaluate_partition_candidates.wgsl:83-297 for one candidate, with
 * the per-invocation avg[]/dir[] arrays of the reference (only invocation p
 * holds partition p's average and direction). */
static void eval_candidate(const or_desc* d, const or_block* b, const or_pinfo* pi, int pc, float* uncor_out,
                           float* samec_out) {
    int tc = d->texel_count;
    uint32_t sums[4][4] = {{0}}, cnt[4] = {0};
    for (int t = 0; t < tc; t++) {
        int p = pi->partition_of_texel[t];
        cnt[p]++;
        for (int c = 0; c < 4; c++) sums[p][c] += wf2u(b->px[t][c]);
    }
    float avg[4][4] = {{0}}; /* avg[p] as held by invocation p */
    int has_avg[4] = {0, 0, 0, 0};
    for (int p = 0; p < pc; p++) {
        float n = (float)cnt[p];
        if (n > 0.0f) {
            for (int c = 0; c < 4; c++) avg[p][c] = (float)sums[p][c] / n;
            has_avg[p] = 1;
        }
    }
    int32_t axis[4][4][4];
    memset(axis, 0, sizeof(axis));
    for (int t = 0; t < tc; t++) {
        int p = pi->partition_of_texel[t];
        float datum[4];
        for (int c = 0; c < 4; c++) datum[c] = b->px[t][c] - ((p == t && t < pc && has_avg[t]) ? avg[t][c] : 0.0f);
        int32_t di[4];
        for (int c = 0; c < 4; c++) di[c] = wf2i(datum[c]);
        for (int ax = 0; ax < 4; ax++)
            if (datum[ax] > 0.0f)
                for (int c = 0; c < 4; c++) axis[p][ax][c] = (int32_t)((uint32_t)axis[p][ax][c] + (uint32_t)di[c]);
    }
    float dir[4][4] = {{0}};
    for (int p = 0; p < pc; p++) {
        float s[4][4];
        for (int ax = 0; ax < 4; ax++)
            for (int c = 0; c < 4; c++) s[ax][c] = (float)axis[p][ax][c];
        int best = 0;
        float mp = dot4(s[0], s[0]);
        for (int ax = 1; ax < 4; ax++) {
            float pr = dot4(s[ax], s[ax]);
            if (pr > mp) { mp = pr; best = ax; }
        }
        normalize4(s[best], dir[p]);
    }
    uint32_t usum = 0, ssum = 0;
    int32_t lmin[4], lmax[4];
    for (int p = 0; p < 4; p++) { lmin[p] = 2147483647; lmax[p] = -2147483647; }
    for (int t = 0; t < tc; t++) {
        int p = pi->partition_of_texel[t];
        const float* px = b->px[t];
        int own = (p == t && t < pc);
        float ud[4], av[4];
        for (int c = 0; c < 4; c++) {
            ud[c] = own ? dir[t][c] : 0.0f;
            av[c] = (own && has_avg[t]) ? avg[t][c] : 0.0f;
        }
        float uparam = dot4(px, ud);
        float dm[4];
        for (int c = 0; c < 4; c++) dm[c] = px[c] - av[c];
        float s = dot4(dm, ud);
        float diff[4];
        for (int c = 0; c < 4; c++) { float proj = av[c] + ud[c] * s; diff[c] = px[c] - proj; diff[c] = diff[c] * diff[c]; }
        float udist = dot4(diff, CW);
        float sd[4];
        normalize4(av, sd);
        float ss = dot4(px, sd);
        for (int c = 0; c < 4; c++) { float proj = sd[c] * ss; diff[c] = px[c] - proj; diff[c] = diff[c] * diff[c]; }
        float sdist = dot4(diff, CW);
        usum += wf2u(udist);
        ssum += wf2u(sdist);
        int32_t pi_ = wf2i(uparam * 4096.0f);
        if (pi_ < lmin[p]) lmin[p] = pi_;
        if (pi_ > lmax[p]) lmax[p] = pi_;
    }
    float tu = (float)usum, ts = (float)ssum;
    for (int p = 0; p < pc; p++) {
        float minp = (float)lmin[p] / 4096.0f, maxp = (float)lmax[p] / 4096.0f;
        float len = wmax(maxp - minp, 1e-7f);
        float tpp = (float)pi->texel_count[p];
        float wie = 0.055f;
        if (tc <= 20) wie = 0.03f;
        else if (tc <= 31) wie = 0.04f;
        else if (tc <= 41) wie = 0.05f;
        wie = wie * wie;
        float ew = tpp * wie;
        float a0[4] = {0, 0, 0, 0};
        if (p == 0 && has_avg[0]) memcpy(a0, avg[0], sizeof(a0));
        tu = tu + ew * (len * len);
        ts = ts + ew * dot4(a0, a0);
    }
    *uncor_out = tu;
    *samec_out = ts;
}

/* pass007_prepare_partitioned_blocks.wgsl:71-174: four candidate positions */
static void select_positions(const float* eu, const float* es, int* pos) {
    int bui[4], bsi[4];
    float bue[4], bse[4];
    for (int i = 0; i < 4; i++) { bui[i] = bsi[i] = 0; bue[i] = bse[i] = 1e30f; }
    for (int i = 0; i < 128; i++) {
        if (eu[i] < bue[3]) {
            bue[3] = eu[i]; bui[3] = i;
            for (int j = 3; j > 0; j--)
                if (bue[j] < bue[j - 1]) { float t = bue[j]; bue[j] = bue[j - 1]; bue[j - 1] = t; int u = bui[j]; bui[j] = bui[j - 1]; bui[j - 1] = u; }
        }
        if (es[i] < bse[3]) {
            bse[3] = es[i]; bsi[3] = i;
            for (int j = 3; j > 0; j--)
                if (bse[j] < bse[j - 1]) { float t = bse[j]; bse[j] = bse[j - 1]; bse[j - 1] = t; int u = bsi[j]; bsi[j] = bsi[j - 1]; bsi[j - 1] = u; }
        }
    }
    int inter[8], emitted = 0;
    for (int i = 0; i < 4; i++) { inter[2 * i] = bui[i]; inter[2 * i + 1] = bsi[i]; }
    for (int i = 0; i < 8 && emitted < 4; i++) {
        int dup = 0;
        for (int j = 0; j < emitted; j++) if (pos[j] == inter[i]) { dup = 1; break; }
        if (!dup) pos[emitted++] = inter[i];
    }
    for (int i = emitted; i < 4; i++) pos[i] = -1;
}

/* Candidate partitionings (packed table positions) for block j of a batch of
 * B blocks whose first block has global index base. */
static void compat_positions(const or_desc* d, const uint8_t* rgba, int w, int h, int bxn, int base, int B, int j,
                             int pc, int* pos) {
    /* partition_ordering / partitioning errors are allocated 920*128 entries but
     * indexed with stride 512 (Q3): blocks j >= 230 see zero errors -> 0,1,2,3 */
    if ((int64_t)j * 512 + 127 >= (int64_t)REF_BATCH * 128) {
        for (int i = 0; i < 4; i++) pos[i] = i;
        return;
    }
    or_block bj, bs;
    uint8_t asg[OR_MAX_TEXELS];
    uint32_t v;
    /* pass004 writes slot block*1024*i: block j reads mismatch(j, 1) (j > 0)
     * or mismatch(B-1, 0) (j == 0), all other selected scores are 0 */
    int src = j == 0 ? B - 1 : j;
    int g = base + src;
    load_block(d, rgba, w, h, g % bxn, g / bxn, &bs);
    kmeans(d, &bs, pc, asg);
    v = mismatch(d, asg, pc, j == 0 ? 0 : 1);
    uint32_t ord[128];
    ordering(d->count_selected[pc - 1], v, ord);
    g = base + j;
    load_block(d, rgba, w, h, g % bxn, g / bxn, &bj);
    float eu[128], es[128];
    for (int c = 0; c < 128; c++)
        eval_candidate(d, &bj, &d->part[(pc - 2) * OR_PARTITIONINGS + ord[c]], pc, &eu[c], &es[c]);
    select_positions(eu, es, pos);
}

/* ---------- exhaustive partition candidates: code/best_partitionings.cpp ---------- */

/* dot_product_squared_difference, best_partitionings.cpp:7-21 */
static float ex_dist(const float* a, const float* b) {
    float d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2], d3 = a[3] - b[3];
    return (((d0 * d0 * CW[0]) + (d1 * d1 * CW[1])) + (d2 * d2 * CW[2])) + (d3 * d3 * CW[3]);
}

/* kmeans_init, best_partitionings.cpp:26-99: first centre texel 145897 % n,
 * then weighted-random picks by cumulative distance against fixed cutoffs */
static void ex_kmeans_init(const or_block* b, int tc, int pc, float centers[4][4]) {
    static const float cutoffs[9] = {0.626220f, 0.932770f, 0.275454f, 0.318558f, 0.240113f,
                                     0.009190f, 0.347661f, 0.731960f, 0.156391f};
    float dist[OR_MAX_TEXELS];
    int n = 0;
    memcpy(centers[n++], b->px[145897 % tc], sizeof(float) * 4);
    float dsum = 0.0f;
    for (int i = 0; i < tc; i++) { dist[i] = ex_dist(b->px[i], centers[0]); dsum += dist[i]; }
    int ci = 3 * (pc - 2);
    for (;;) {
        float summa = 0.0f, cut = dsum * cutoffs[ci++];
        int next = 0;
        for (; next < tc; next++) { summa += dist[next]; if (summa >= cut) break; }
        if (next > tc - 1) next = tc - 1;
        memcpy(centers[n++], b->px[next], sizeof(float) * 4);
        if (n >= pc) break;
        dsum = 0.0f;
        for (int i = 0; i < tc; i++) {
            float d = ex_dist(b->px[i], centers[n - 1]);
            dist[i] = d < dist[i] ? d : dist[i];
            dsum += dist[i];
        }
    }
}

/* kmeans_assign, best_partitionings.cpp:101-151 (with its empty-cluster fix) */
static void ex_kmeans_assign(const or_block* b, int tc, int pc, float centers[4][4], uint8_t* part) {
    int cnt[4] = {0, 0, 0, 0};
    for (int i = 0; i < tc; i++) {
        float bd = 3.402823466e38f;
        int bp = 0;
        for (int j = 0; j < pc; j++) {
            float d = ex_dist(b->px[i], centers[j]);
            if (d < bd) { bd = d; bp = j; }
        }
        part[i] = (uint8_t)bp;
        cnt[bp]++;
    }
    int problem;
    do {
        problem = 0;
        for (int i = 0; i < pc; i++)
            if (cnt[i] == 0) {
                problem = 1;
                cnt[part[i]]--;
                cnt[i]++;
                part[i] = (uint8_t)i;
            }
    } while (problem);
}

/* kmeans_update, best_partitionings.cpp:153-194 */
static void ex_kmeans_update(const or_block* b, int tc, int pc, float centers[4][4], const uint8_t* part) {
    float sum[4][4] = {{0}};
    int cnt[4] = {0, 0, 0, 0};
    for (int i = 0; i < tc; i++) {
        for (int c = 0; c < 4; c++) sum[part[i]][c] += b->px[i][c];
        cnt[part[i]]++;
    }
    for (int i = 0; i < pc; i++) {
        if (cnt[i] == 0) { for (int c = 0; c < 4; c++) centers[i][c] = 0.0f; continue; }
        float sc = 1.0f / (float)cnt[i];
        for (int c = 0; c < 4; c++) centers[i][c] = sum[i][c] * sc;
    }
}

/* compute_kmeans_partition_ordering + get_partition_ordering_by_mismatch_bits,
 * best_partitionings.cpp:364-397, 399-447: 3 k-means passes, bitmaps over
 * the first min(texels, 64) k-means texels, mismatch per selected
 * partitioning, stable counting sort by mismatch.  Returns the count. */
static int ex_ordering(const or_desc* d, const or_block* b, int pc, uint16_t* order) {
    int tc = d->texel_count;
    float centers[4][4];
    uint8_t part[OR_MAX_TEXELS];
    for (int it = 0; it < 3; it++) {
        if (it == 0) ex_kmeans_init(b, tc, pc, centers);
        else ex_kmeans_update(b, tc, pc, centers, part);
        ex_kmeans_assign(b, tc, pc, centers, part);
    }
    int nk = tc < 64 ? tc : 64;
    uint64_t a[4] = {0, 0, 0, 0};
    for (int i = 0; i < nk; i++) a[part[d->kmeans_texels[i]]] |= 1ull << i;
    int nsel = d->count_selected[pc - 1];
    uint8_t mis[OR_PARTITIONINGS];
    for (int i = 0; i < nsel; i++) {
        const uint64_t* bb = d->coverage[pc - 2][i];
#define P(x, y) popc64(a[x] ^ bb[y])
#define MIN(x, y) ((x) < (y) ? (x) : (y))
        int v;
        if (pc == 2) v = MIN(P(0, 0) + P(1, 1), P(0, 1) + P(1, 0));
        else if (pc == 3) {
            int v0 = MIN(P(1, 1) + P(2, 2), P(1, 2) + P(2, 1)) + P(0, 0);
            int v1 = MIN(P(1, 0) + P(2, 2), P(1, 2) + P(2, 0)) + P(0, 1);
            int v2 = MIN(P(1, 0) + P(2, 1), P(1, 1) + P(2, 0)) + P(0, 2);
            v = MIN(v0, MIN(v1, v2));
        } else {
            int mx23 = MIN(P(2, 2) + P(3, 3), P(2, 3) + P(3, 2)), mx13 = MIN(P(2, 1) + P(3, 3), P(2, 3) + P(3, 1));
            int mx12 = MIN(P(2, 1) + P(3, 2), P(2, 2) + P(3, 1)), mx03 = MIN(P(2, 0) + P(3, 3), P(2, 3) + P(3, 0));
            int mx02 = MIN(P(2, 0) + P(3, 2), P(2, 2) + P(3, 0)), mx01 = MIN(P(2, 1) + P(3, 0), P(2, 0) + P(3, 1));
            int v0 = P(0, 0) + MIN(P(1, 1) + mx23, MIN(P(1, 2) + mx13, P(1, 3) + mx12));
            int v1 = P(0, 1) + MIN(P(1, 0) + mx23, MIN(P(1, 2) + mx03, P(1, 3) + mx02));
            int v2 = P(0, 2) + MIN(P(1, 1) + mx03, MIN(P(1, 0) + mx13, P(1, 3) + mx01));
            int v3 = P(0, 3) + MIN(P(1, 1) + mx02, MIN(P(1, 2) + mx01, P(1, 0) + mx12));
            v = MIN(v0, MIN(v1, MIN(v2, v3)));
        }
#undef P
#undef MIN
        mis[i] = (uint8_t)(v / 2);
    }
    uint16_t msc[64];
    memset(msc, 0, sizeof(msc));
    for (int i = 0; i < nsel; i++) msc[mis[i]]++;
    uint16_t sum = 0;
    for (int i = 0; i < nk; i++) { uint16_t c = msc[i]; msc[i] = sum; sum = (uint16_t)(sum + c); }
    /* a mismatch equal to the texel count is past the running sum, so its
     * entries can land past the sequence: those writes are dropped (the
     * reference writes past its array there) */
    for (int i = 0; i < nsel; i++) {
        int pos = msc[mis[i]]++;
        if (pos < nsel) order[pos] = (uint16_t)i;
    }
    return nsel;
}

/* normalize, averages_and_directions.cpp:11-22 (zero vector -> zero) */
static void ex_normalize(const float* v, float* o) {
    float n = ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3];
    if (n != 0.0f) { n = sqrtf(n); for (int c = 0; c < 4; c++) o[c] = v[c] / n; }
    else for (int c = 0; c < 4; c++) o[c] = 0.0f;
}

/* insert_result, best_partitionings.cpp:460-496 */
static void ex_insert(float e, int id, float* be, int* bi) {
    if (e >= be[3]) return;
    for (int i = 0; i < 4; i++) {
        if (e > be[i]) continue;
        for (int j = 3; j > i; j--) { be[j] = be[j - 1]; bi[j] = bi[j - 1]; }
        be[i] = e;
        bi[i] = id;
        break;
    }
}

/* find_best_partition_candidates, best_partitionings.cpp:498-737, over every
 * selected partitioning (search limit = all), 4 requested candidates, with
 * the line fits of averages_and_directions.cpp (4 components when alpha
 * varies, RGB otherwise).  Writes 4 packed table entries (-1 = none). */
/* The line fits of find_best_partition_candidates (best_partitionings.cpp:
 * 520-737) over the packed table entries seq[0 .. n), in that order, and
 * the 4 best of each fit kept by insert_result, interleaved and
 * de-duplicated into pos (-1 = none). */
static void line_fit_positions(const or_desc* d, const or_block* b, int pc, const uint16_t* seq, int n, int* pos) {
    int tc = d->texel_count;
    float wie = 0.055f;
    if (tc <= 20) wie = 0.03f;
    else if (tc <= 31) wie = 0.04f;
    else if (tc <= 41) wie = 0.05f;
    wie = wie * wie;
    int req = n < 4 ? n : 4;
    float ube[4], sbe[4];
    int ubi[4] = {0, 0, 0, 0}, sbi[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; i++) ube[i] = sbe[i] = 1e30f;
    int rgba = !b->constant_alpha;
    for (int s = 0; s < n; s++) {
        const or_pinfo* pi = &d->part[(pc - 2) * OR_PARTITIONINGS + seq[s]];
        float sum[4][4] = {{0}};
        for (int i = 0; i < tc; i++)
            for (int c = 0; c < 4; c++) sum[pi->partition_of_texel[i]][c] += b->px[i][c];
        float avg[4][4], bdir[4][4], bavg[4][4], amod[4][4], len[4];
        for (int p = 0; p < pc; p++) {
            for (int c = 0; c < 4; c++) avg[p][c] = sum[p][c] / (float)pi->texel_count[p];
            float sx[4] = {0}, sy[4] = {0}, sz[4] = {0}, sw[4] = {0};
            for (int i = 0; i < tc; i++) {
                if (pi->partition_of_texel[i] != p) continue;
                float t[4];
                for (int c = 0; c < 4; c++) t[c] = b->px[i][c] - avg[p][c];
                if (t[0] > 0) for (int c = 0; c < 4; c++) sx[c] += t[c];
                if (t[1] > 0) for (int c = 0; c < 4; c++) sy[c] += t[c];
                if (t[2] > 0) for (int c = 0; c < 4; c++) sz[c] += t[c];
                if (rgba && t[3] > 0) for (int c = 0; c < 4; c++) sw[c] += t[c];
            }
            float* best = sx;
            float bs, pr;
            if (rgba) {
                bs = dot4(sx, sx);
                pr = dot4(sy, sy); if (pr > bs) { bs = pr; best = sy; }
                pr = dot4(sz, sz); if (pr > bs) { bs = pr; best = sz; }
                pr = dot4(sw, sw); if (pr > bs) { bs = pr; best = sw; }
            } else {
                bs = dot3(sx, sx);
                pr = dot3(sy, sy); if (pr > bs) { bs = pr; best = sy; }
                pr = dot3(sz, sz); if (pr > bs) { bs = pr; best = sz; }
            }
            float dir[4] = {best[0], best[1], best[2], rgba ? best[3] : 0.0f};
            ex_normalize(avg[p], bavg[p]);
            ex_normalize(dir, bdir[p]);
            float ud = rgba ? dot4(avg[p], bdir[p]) : dot3(avg[p], bdir[p]);
            for (int c = 0; c < 4; c++) amod[p][c] = avg[p][c] - bdir[p][c] * ud;
        }
        float ue = 0.0f, se = 0.0f;
        int nch = rgba ? 4 : 3;
        for (int p = 0; p < pc; p++) {
            float lo = 1e10f, hi = -1e10f;
            for (int i = 0; i < tc; i++) {
                if (pi->partition_of_texel[i] != p) continue;
                const float* px = b->px[i];
                float up = px[0] * bdir[p][0] + px[1] * bdir[p][1] + px[2] * bdir[p][2];
                if (rgba) up = up + px[3] * bdir[p][3];
                lo = up < lo ? up : lo;
                hi = up > hi ? up : hi;
                float err = 0.0f, sp = px[0] * bavg[p][0] + px[1] * bavg[p][1] + px[2] * bavg[p][2];
                if (rgba) sp = sp + px[3] * bavg[p][3];
                float serr = 0.0f;
                for (int c = 0; c < nch; c++) {
                    float dd = (amod[p][c] - px[c]) + (up * bdir[p][c]);
                    float ds = sp * bavg[p][c] - px[c];
                    err = c ? err + CW[c] * dd * dd : CW[c] * dd * dd;
                    serr = c ? serr + CW[c] * ds * ds : CW[c] * ds * ds;
                }
                ue += err;
                se += serr;
            }
            float l = hi - lo;
            len[p] = l > 1e-7f ? l : 1e-7f;
        }
        for (int p = 0; p < pc; p++) {
            float ew = (float)pi->texel_count[p] * wie;
            float uv[4], sv[4];
            for (int c = 0; c < 4; c++) { uv[c] = bdir[p][c] * len[p]; sv[c] = bavg[p][c] * len[p]; }
            ue += (rgba ? dot4(uv, uv) : dot3(uv, uv)) * ew;
            se += (rgba ? dot4(sv, sv) : dot3(sv, sv)) * ew;
        }
        ex_insert(ue, seq[s], ube, ubi);
        ex_insert(se, seq[s], sbe, sbi);
    }
    int inter[8], emitted = 0;
    for (int i = 0; i < req; i++) { inter[2 * i] = ubi[i]; inter[2 * i + 1] = sbi[i]; }
    for (int i = 0; i < 4; i++) pos[i] = -1;
    for (int i = 0; i < 2 * req && emitted < req; i++) {
        int dup = 0;
        for (int j = 0; j < emitted; j++) if (pos[j] == inter[i]) dup = 1;
        if (!dup) pos[emitted++] = inter[i];
    }
}

static void exhaustive_positions(const or_desc* d, const or_block* b, int pc, int limit, int* pos) {
    /* positions the counting sort leaves unwritten (its running sum stops
     * before the largest possible mismatch count) read 0 here; the
     * reference reads an uninitialised stack array there */
    uint16_t seq[OR_PARTITIONINGS];
    memset(seq, 0, sizeof(seq));
    int n = ex_ordering(d, b, pc, seq);
    if (limit > 0 && limit < n) n = limit;  /* partition_search_limit, best_partitionings.cpp:520 */
    line_fit_positions(d, b, pc, seq, n, pos);
}

/* Candidates from a list of partition seeds (a predictor's most likely
 * seeds, best first): the seeds that are selected partitionings of this
 * count, in list order; more than 4 seeds are ranked by the line fits of
 * line_fit_positions (after the first mis_n partitionings of the mismatch
 * order, seeds among those skipped), up to 4 are taken as they are.  No
 * counterpart in the reference: the learned policy's candidates. */
static void seeded_positions(const or_desc* d, const or_block* b, int pc, const int32_t* seeds, int nseeds,
                             int mis_n, int* pos) {
    uint16_t seq[2 * OR_PARTITIONINGS];
    int n = 0;
    if (nseeds > 4 && mis_n > 0) {
        /* the first mis_n partitionings of the exhaustive search's mismatch
         * order come first (every slot of that order is written: a
         * mismatch count stays below the k-means texel count) */
        for (int i = 0; i < OR_PARTITIONINGS; i++) seq[i] = 0;
        int nsel = ex_ordering(d, b, pc, seq);
        n = mis_n < nsel ? mis_n : nsel;
    }
    int n_order = n;
    for (int j = 0; j < nseeds; j++) {
        int seed = seeds[j];
        if (seed < 0 || seed >= OR_PARTITIONINGS) break;
        for (int e = 0; e < d->count_selected[pc - 1]; e++)
            if (d->part[(pc - 2) * OR_PARTITIONINGS + e].partition_index == seed) {
                int dup = 0;
                for (int t = 0; t < n_order; t++) dup |= seq[t] == e;
                if (!dup) seq[n++] = (uint16_t)e;
                break;
            }
    }
    for (int i = 0; i < 4; i++) pos[i] = -1;
    if (nseeds > 4) {
        line_fit_positions(d, b, pc, seq, n, pos);
    } else {
        for (int i = 0; i < n && i < 4; i++) pos[i] = seq[i];
    }
}

int oracle_debug_exhaustive_candidates(const uint8_t* rgba, int width, int height, int bx, int by, int block_index,
                                       int partition_count, int32_t* entries_out) {
    or_desc* d = or_desc_get(bx, by);
    if (!d || partition_count < 2 || partition_count > 4) return -1;
    int bxn = (width + bx - 1) / bx;
    or_block blk;
    load_block(d, rgba, width, height, block_index % bxn, block_index / bxn, &blk);
    int pos[4];
    exhaustive_positions(d, &blk, partition_count, 0, pos);
    for (int i = 0; i < 4; i++) entries_out[i] = pos[i];
    return 0;
}

int oracle_debug_seeded_candidates(const uint8_t* rgba, int width, int height, int bx, int by, int block_index,
                                   int partition_count, const int32_t* seeds, int nseeds, int mismatch_candidates,
                                   int32_t* entries_out) {
    or_desc* d = or_desc_get(bx, by);
    if (!d || partition_count < 2 || partition_count > 4 || nseeds < 0 || nseeds > OR_PARTITIONINGS) return -1;
    int bxn = (width + bx - 1) / bx;
    or_block blk;
    load_block(d, rgba, width, height, block_index % bxn, block_index / bxn, &blk);
    int pos[4];
    seeded_positions(d, &blk, partition_count, seeds, nseeds, mismatch_candidates, pos);
    for (int i = 0; i < 4; i++) entries_out[i] = pos[i];
    return 0;
}

int oracle_debug_kmeans_candidates(const uint8_t* rgba, int width, int height, int bx, int by, int block_index,
                                   int partition_count, int batch_size, int32_t* positions_out) {
    or_desc* d = or_desc_get(bx, by);
    if (!d || partition_count < 2 || partition_count > 4) return -1;
    int bxn = (width + bx - 1) / bx, byn = (height + by - 1) / by;
    int nb = bxn * byn;
    if (batch_size <= 0) batch_size = REF_BATCH;
    int base = (block_index / batch_size) * batch_size;
    int B = nb - base < batch_size ? nb - base : batch_size;
    int pos[4];
    compat_positions(d, rgba, width, height, bxn, base, B, block_index - base, partition_count, pos);
    for (int i = 0; i < 4; i++) positions_out[i] = pos[i];
    return 0;
}

int oracle_encode(const uint8_t* rgba, int width, int height, int bx, int by, const or_params* prm,
                  uint8_t* out_blocks, or_symbolic* out_sym, or_stats* stats) {
    or_desc* d = or_desc_get(bx, by);
    if (!d || !rgba || width <= 0 || height <= 0) return -1;
    int maxp = prm && prm->max_partitions > 0 ? prm->max_partitions : 4;
    int batch = prm && prm->batch_size > 0 ? prm->batch_size : REF_BATCH;
    int threads = prm && prm->threads > 0 ? prm->threads : 1;
    int policy = prm ? prm->policy : 0;
    int search_limit = prm ? prm->search_limit : 0;
    int bxn = (width + bx - 1) / bx, byn = (height + by - 1) / by;
    int nb = bxn * byn;
    int first = prm ? prm->first_block : 0;
    int count = prm && prm->block_count > 0 ? prm->block_count : nb - first;
    if (first < 0 || first + count > nb) return -2;
    or_stats st = {0, 0, 0};
    uint64_t nshort = 0, npb = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4) reduction(+ : nshort, npb)
    for (int k = 0; k < count; k++) {
        int g = first + k;
        int base = (g / batch) * batch;
        int B = nb - base < batch ? nb - base : batch;
        int j = g - base;
        or_block blk;
        load_block(d, rgba, width, height, g % bxn, g / bxn, &blk);
        or_symbolic best;
        memset(&best, 0, sizeof(best));
        best.errorval = 1e30f;
        or_stats ls = {0, 0, 0};
        for (int pc = 1; pc <= maxp; pc++) {
            or_symbolic cur;
            memset(&cur, 0, sizeof(cur));
            cur.errorval = OR_ERR_DEFAULT;
            int have = 0;
            int pos[4] = {0, -1, -1, -1};
            int npos = 1;
            if (pc > 1) {
                if (policy == 2) exhaustive_positions(d, &blk, pc, search_limit, pos);
                else if (policy == 3)
                    seeded_positions(d, &blk, pc, prm->seed_list + (pc - 2) * prm->seed_list_len, prm->seed_list_len,
                                     prm->mismatch_candidates, pos);
                else compat_positions(d, rgba, width, height, bxn, base, B, j, pc, pos);
                npos = 4;
            }
            for (int c = 0; c < npos; c++) {
                if (pos[c] < 0) continue;
                or_block pbk = blk;
                if (pc > 1) {
                    const or_pinfo* pi = &d->part[(pc - 2) * OR_PARTITIONINGS + pos[c]];
                    memcpy(pbk.tp, pi->partition_of_texel, sizeof(pbk.tp));
                    for (int i = 0; i < 4; i++) pbk.ppc[i] = pi->texel_count[i];
                    pbk.partitioning_idx = pi->partition_index;
                }
                or_symbolic s;
                int pb_in_batch = pc == 1 ? j : j * 4 + c;
                search_pb(d, &pbk, pc, pb_in_batch, &s, &ls);
                npb++;
                /* compression_pipeline.cpp:507-512: strict < keeps the first */
                if (!have || s.errorval < cur.errorval) { cur = s; have = 1; }
            }
            if (have && cur.errorval < best.errorval) best = cur;
        }
        nshort += ls.short_candidate_lists;
        if (out_sym) out_sym[k] = best;
        if (out_blocks) oracle_symbolic_to_physical(bx, by, &best, out_blocks + (size_t)k * 16);
    }
    st.blocks = (uint64_t)count;
    st.partitioned_blocks = npb;
    st.short_candidate_lists = nshort;
    if (stats) *stats = st;
    return 0;
}

/* ====================================================================== */
/* Stage-level exports for the oracle's own unit tests (tests/           */
/* test_oracle_pins.py): one pass at a time on explicit inputs, so that   */
/* each pass can be checked against closed forms, float64 recomputations */
/* and brute force.  Texel values are in the search's scale (u8/255 *     */
/* 65536); tp gives each texel's partition.                               */
/* ====================================================================== */

static int dbg_block(const or_desc* d, const float* px, const uint8_t* tp, int pc, or_block* b) {
    memset(b, 0, sizeof(*b));
    float amin = 1e38f, amax = 0.0f;
    int gray = 1;
    for (int i = 0; i < d->texel_count; i++) {
        for (int c = 0; c < 4; c++) b->px[i][c] = px[i * 4 + c];
        b->tp[i] = tp ? tp[i] : 0;
        if (b->tp[i] >= pc) return -1;
        b->ppc[b->tp[i]]++;
        amax = fmaxf(amax, b->px[i][3]);
        amin = fminf(amin, b->px[i][3]);
        gray = gray && (b->px[i][0] == b->px[i][1] && b->px[i][0] == b->px[i][2]);
    }
    b->grayscale = gray;
    b->constant_alpha = amax == amin;
    return 0;
}

static const or_dm* dbg_dm(const or_desc* d, int m) {
    if (m < 0 || m >= d->ndm) return NULL;
    return &d->dm[d->valid_dm[m]];
}

/* pass01: out = avg[16], dir[16], ep0[16], ep1[16], weights[tc], wes[tc],
 * is_const_wes, min_weight_cutoff.  Returns the texel count. */
int oracle_debug_pass01(int bx, int by, const float* px, const uint8_t* tp, int pc, float* out) {
    or_desc* d = or_desc_get(bx, by);
    or_block b;
    if (!d || dbg_block(d, px, tp, pc, &b)) return -1;
    or_ideal o;
    pass01(d, &b, &o);
    int tc = d->texel_count, n = 0;
    for (int p = 0; p < 4; p++) for (int c = 0; c < 4; c++) out[n++] = o.part[p].avg[c];
    for (int p = 0; p < 4; p++) for (int c = 0; c < 4; c++) out[n++] = o.part[p].dir[c];
    for (int p = 0; p < 4; p++) for (int c = 0; c < 4; c++) out[n++] = o.part[p].ep0[c];
    for (int p = 0; p < 4; p++) for (int c = 0; c < 4; c++) out[n++] = o.part[p].ep1[c];
    for (int i = 0; i < tc; i++) out[n++] = o.weights[i];
    for (int i = 0; i < tc; i++) out[n++] = o.wes[i];
    out[n++] = (float)o.is_const_wes;
    out[n++] = o.min_weight_cutoff;
    return tc;
}

static void dbg_ideal(const or_desc* d, const float* weights, const float* wes, int is_const, or_ideal* o) {
    memset(o, 0, sizeof(*o));
    for (int i = 0; i < d->texel_count; i++) { o->weights[i] = weights[i]; o->wes[i] = wes[i]; }
    o->is_const_wes = is_const;
}

/* pass02 for trial decimation mode m: decimated weights; returns the weight count. */
int oracle_debug_pass02(int bx, int by, int m, const float* weights, const float* wes, int is_const, float* dw) {
    or_desc* d = or_desc_get(bx, by);
    const or_dm* di = d ? dbg_dm(d, m) : NULL;
    if (!di) return -1;
    or_ideal o;
    dbg_ideal(d, weights, wes, is_const, &o);
    pass02(d, di, &o, dw);
    return di->weight_count;
}

/* pass03: 16 angular offsets of trial decimation mode m. */
int oracle_debug_pass03(int bx, int by, int m, const float* dw, float* ao) {
    or_desc* d = or_desc_get(bx, by);
    const or_dm* di = d ? dbg_dm(d, m) : NULL;
    if (!di) return -1;
    pass03(d, di, dw, ao);
    return 16;
}

/* pass04: per angular step {lowest, span, error, cut_low, cut_high};
 * returns max_angular_steps (its max_quant_steps in info[0], its
 * max_quant_level in info[1]). */
int oracle_debug_pass04(int bx, int by, int m, const float* dw, const float* ao, float* hl5, int* info) {
    or_desc* d = or_desc_get(bx, by);
    const or_dm* di = d ? dbg_dm(d, m) : NULL;
    if (!di) return -1;
    or_hl hl[16];
    pass04(di, dw, ao, hl);
    for (int s = 0; s < di->max_angular_steps; s++) {
        hl5[s * 5 + 0] = hl[s].lowest;
        hl5[s * 5 + 1] = (float)hl[s].span;
        hl5[s * 5 + 2] = hl[s].error;
        hl5[s * 5 + 3] = hl[s].cut_low;
        hl5[s * 5 + 4] = hl[s].cut_high;
    }
    if (info) { info[0] = di->max_quant_steps; info[1] = di->max_quant_level; }
    return di->max_angular_steps;
}

/* pass05: low/high weight values per quant level 0..max_quant_level from
 * pass04's records (same layout as oracle_debug_pass04). */
int oracle_debug_pass05(int bx, int by, int m, const float* hl5, const float* ao, float* low, float* high) {
    or_desc* d = or_desc_get(bx, by);
    const or_dm* di = d ? dbg_dm(d, m) : NULL;
    if (!di) return -1;
    or_hl hl[16];
    for (int s = 0; s < di->max_angular_steps; s++) {
        hl[s].lowest = hl5[s * 5 + 0];
        hl[s].span = (int)hl5[s * 5 + 1];
        hl[s].error = hl5[s * 5 + 2];
        hl[s].cut_low = hl5[s * 5 + 3];
        hl[s].cut_high = hl5[s * 5 + 4];
    }
    pass05(di, hl, ao, low, high);
    return di->max_quant_level;
}

/* Trial block mode k: {decimation trial index, weight quant, weight bits,
 * weight count, raw mode}. */
int oracle_debug_trial_mode(int bx, int by, int k, int* info) {
    or_desc* d = or_desc_get(bx, by);
    if (!d || k < 0 || k >= d->nbm) return -1;
    const or_bm* bm = &d->bm[d->valid_bm[k]];
    info[0] = d->valid_bm_dml[k];
    info[1] = bm->quant_mode;
    info[2] = bm->weight_bits;
    info[3] = d->dm[bm->decimation_mode].weight_count;
    info[4] = bm->mode_index;
    return d->nbm;
}

/* pass07 for trial block mode k (dw: its decimation mode's weights):
 * error, remaining bit count, quantised weights (0..64, whole groups of four
 * kept).  Returns the weight count. */
int oracle_debug_pass07(int bx, int by, int k, int pc, const float* weights, const float* wes, int is_const,
                        float cutoff, const float* dw, float low, float high, float* error, int* bitcount,
                        uint8_t* qw) {
    or_desc* d = or_desc_get(bx, by);
    if (!d || k < 0 || k >= d->nbm || pc < 1 || pc > 4) return -1;
    const or_bm* bm = &d->bm[d->valid_bm[k]];
    or_ideal o;
    dbg_ideal(d, weights, wes, is_const, &o);
    o.min_weight_cutoff = cutoff;
    or_qres r;
    pass07(d, bm, &d->dm[bm->decimation_mode], &o, dw, low, high, pc, &r);
    *error = r.error;
    *bitcount = r.bitcount;
    memcpy(qw, r.qw, OR_MAX_WEIGHTS);
    return d->dm[bm->decimation_mode].weight_count;
}

/* pass08 (after pass01 on the same block): per partition {rgb_scale_error,
 * rgb_luma_error, luminance_error, alpha_drop_error, can_offset,
 * can_blue_contract}. */
int oracle_debug_pass08(int bx, int by, const float* px, const uint8_t* tp, int pc, float* out) {
    or_desc* d = or_desc_get(bx, by);
    or_block b;
    if (!d || dbg_block(d, px, tp, pc, &b)) return -1;
    or_ideal o;
    pass01(d, &b, &o);
    or_ece e[4];
    pass08(d, &b, &o, pc, e);
    for (int p = 0; p < pc; p++) {
        out[p * 6 + 0] = e[p].rgb_scale_error;
        out[p * 6 + 1] = e[p].rgb_luma_error;
        out[p * 6 + 2] = e[p].luminance_error;
        out[p * 6 + 3] = e[p].alpha_drop_error;
        out[p * 6 + 4] = (float)e[p].can_offset;
        out[p * 6 + 5] = (float)e[p].can_blue_contract;
    }
    return pc;
}

/* pass09 from explicit inputs: endpoints ep0/ep1 [4][4], texel counts per
 * partition, pass08 records (6 floats per partition as above).  Out:
 * err[4][21][4], fmt[4][21][4]. */
int oracle_debug_pass09(int pc, const float* ep0, const float* ep1, const int* counts, const float* ece6, float* err,
                        int* fmt) {
    if (pc < 1 || pc > 4) return -1;
    or_block b;
    memset(&b, 0, sizeof(b));
    or_ideal o;
    memset(&o, 0, sizeof(o));
    or_ece e[4];
    for (int p = 0; p < pc; p++) {
        b.ppc[p] = counts[p];
        for (int c = 0; c < 4; c++) { o.part[p].ep0[c] = ep0[p * 4 + c]; o.part[p].ep1[c] = ep1[p * 4 + c]; }
        e[p].rgb_scale_error = ece6[p * 6 + 0];
        e[p].rgb_luma_error = ece6[p * 6 + 1];
        e[p].luminance_error = ece6[p * 6 + 2];
        e[p].alpha_drop_error = ece6[p * 6 + 3];
        e[p].can_offset = ece6[p * 6 + 4] != 0.0f;
        e[p].can_blue_contract = ece6[p * 6 + 5] != 0.0f;
    }
    float cerr[4][21][4];
    int cfmt[4][21][4];
    memset(cerr, 0, sizeof(cerr));
    memset(cfmt, 0, sizeof(cfmt));
    pass09(&b, &o, e, pc, cerr, cfmt);
    memcpy(err, cerr, sizeof(cerr));
    memcpy(fmt, cfmt, sizeof(cfmt));
    return pc;
}

/* pass10 from pass09's tables: comb_err[21][13], comb_fmt[21][13][4]. */
int oracle_debug_pass10(int pc, const float* err, const int* fmt, float* comb_err, int* comb_fmt) {
    if (pc < 2 || pc > 4) return -1;
    float cerr[4][21][4];
    int cfmt[4][21][4];
    memcpy(cerr, err, sizeof(cerr));
    memcpy(cfmt, fmt, sizeof(cfmt));
    static __thread or_comb comb[21][13];
    pass10(pc, cerr, cfmt, comb);
    for (int q = 0; q < 21; q++)
        for (int k = 0; k < 13; k++) {
            comb_err[q * 13 + k] = comb[q][k].error;
            for (int p = 0; p < 4; p++) comb_fmt[(q * 13 + k) * 4 + p] = comb[q][k].formats[p];
        }
    return 0;
}

/* pass11 from explicit inputs: weight error and remaining bits of a block
 * mode, pass09 tables (1 partition) or pass10 tables (2..4).  Out: total
 * error, colour quant level, its modified level, formats[4]. */
int oracle_debug_pass11(int pc, float weight_error, int bitcount, const float* err, const int* fmt,
                        const float* comb_err, const int* comb_fmt, float* total, int* ql, int* ql_mod, int* formats) {
    if (pc < 1 || pc > 4) return -1;
    float cerr[4][21][4];
    int cfmt[4][21][4];
    memcpy(cerr, err, sizeof(cerr));
    memcpy(cfmt, fmt, sizeof(cfmt));
    static __thread or_comb comb[21][13];
    for (int q = 0; q < 21; q++)
        for (int k = 0; k < 13; k++) {
            comb[q][k].error = comb_err ? comb_err[q * 13 + k] : OR_ERR_DEFAULT;
            for (int p = 0; p < 4; p++) comb[q][k].formats[p] = comb_fmt ? comb_fmt[(q * 13 + k) * 4 + p] : 0;
        }
    or_qres qr;
    memset(&qr, 0, sizeof(qr));
    qr.error = weight_error;
    qr.bitcount = bitcount;
    or_ccr c;
    memset(&c, 0, sizeof(c));
    pass11(pc, &qr, cerr, cfmt, comb, &c);
    *total = c.total_error;
    *ql = c.ql;
    *ql_mod = c.ql_mod;
    for (int p = 0; p < 4; p++) formats[p] = c.formats[p];
    return 0;
}

/* pass12: indices of the 8 best of n errors (insertion, strict <), and
 * their errors. */
int oracle_debug_pass12(const float* errors, int n, int* idx, float* top) {
    float te[8];
    int ti[8];
    for (int i = 0; i < 8; i++) { te[i] = OR_ERR_DEFAULT; ti[i] = 0; }
    for (int k = 0; k < n; k++) {
        float e = errors[k];
        if (e < te[7]) {
            te[7] = e; ti[7] = k;
            for (int j = 7; j > 0; j--)
                if (te[j] < te[j - 1]) {
                    float fe = te[j]; te[j] = te[j - 1]; te[j - 1] = fe;
                    int fi = ti[j]; ti[j] = ti[j - 1]; ti[j - 1] = fi;
                }
        }
    }
    for (int i = 0; i < 8; i++) { idx[i] = ti[i]; top[i] = te[i]; }
    return 0;
}

/* pass13 for trial block mode k of a block with quantised weights qw
 * (0..64): endpoints in/out ep[2][4][4] (ep0 then ep1), averages avg[4][4]
 * (from pass01), rgbs[4][4] out. */
int oracle_debug_pass13(int bx, int by, int k, const float* px, const uint8_t* tp, int pc, const float* avg,
                        float* ep, const uint8_t* qw, float* rgbs) {
    or_desc* d = or_desc_get(bx, by);
    or_block b;
    if (!d || k < 0 || k >= d->nbm || dbg_block(d, px, tp, pc, &b)) return -1;
    or_cand c;
    memset(&c, 0, sizeof(c));
    c.bm_packed = d->valid_bm[k];
    for (int p = 0; p < 4; p++)
        for (int i = 0; i < 4; i++) {
            c.part[p].avg[i] = avg[p * 4 + i];
            c.part[p].ep0[i] = ep[p * 4 + i];
            c.part[p].ep1[i] = ep[16 + p * 4 + i];
        }
    for (int i = 0; i < OR_MAX_WEIGHTS; i++) c.qw[i] = qw[i];
    float r[4][4];
    memset(r, 0, sizeof(r));
    pass13(d, &b, pc, &c, r);
    for (int p = 0; p < 4; p++)
        for (int i = 0; i < 4; i++) {
            ep[p * 4 + i] = c.part[p].ep0[i];
            ep[16 + p * 4 + i] = c.part[p].ep1[i];
            rgbs[p * 4 + i] = r[p][i];
        }
    return 0;
}

/* pass14's per-partition endpoint packer (pack_color_endpoints): packed
 * values (0..255 unquantised), returns the format it chose. */
int oracle_debug_pack_endpoints(const float* ep0, const float* ep1, const float* rgbs, int format, int q,
                                uint32_t* packed) {
    return or_pack_color_endpoints(ep0, ep1, rgbs, format, q, packed);
}

/* pass15's unpacker: endpoints 0..255 (before the x257 of pass15). */
int oracle_debug_unpack_endpoints(int format, const uint32_t* packed, int* e0, int* e1) {
    or_unpack_color_endpoints(format, packed, e0, e1);
    return 0;
}

/* pass16 for trial block mode k: qw (0..64) realigned in place against the
 * unpacked endpoints ue0/ue1 [4][4] (x257 scale). */
int oracle_debug_pass16(int bx, int by, int k, const float* px, const uint8_t* tp, int pc, const int* ue0,
                        const int* ue1, uint8_t* qw) {
    or_desc* d = or_desc_get(bx, by);
    or_block b;
    if (!d || k < 0 || k >= d->nbm || dbg_block(d, px, tp, pc, &b)) return -1;
    or_cand c;
    memset(&c, 0, sizeof(c));
    c.bm_packed = d->valid_bm[k];
    for (int i = 0; i < OR_MAX_WEIGHTS; i++) c.qw[i] = qw[i];
    int u0[4][4], u1[4][4];
    memcpy(u0, ue0, sizeof(u0));
    memcpy(u1, ue1, sizeof(u1));
    pass16(d, &b, pc, &c, u0, u1);
    for (int i = 0; i < OR_MAX_WEIGHTS; i++) qw[i] = (uint8_t)c.qw[i];
    return 0;
}

/* pass17 for trial block mode k: the block error of weights qw and
 * endpoints ue0/ue1 (x257 scale). */
float oracle_debug_pass17(int bx, int by, int k, const float* px, const uint8_t* tp, int pc, const int* ue0,
                          const int* ue1, const uint8_t* qw) {
    or_desc* d = or_desc_get(bx, by);
    or_block b;
    if (!d || k < 0 || k >= d->nbm || dbg_block(d, px, tp, pc, &b)) return -1.0f;
    or_cand c;
    memset(&c, 0, sizeof(c));
    c.bm_packed = d->valid_bm[k];
    for (int i = 0; i < OR_MAX_WEIGHTS; i++) c.qw[i] = qw[i];
    int u0[4][4], u1[4][4];
    memcpy(u0, ue0, sizeof(u0));
    memcpy(u1, ue1, sizeof(u1));
    return pass17(d, &b, pc, &c, u0, u1);
}

/* Full per-partitioned-block search (passes 01-18) on an explicit block:
 * the chosen symbolic block.  pb_in_batch places it in the reference's
 * batch (pass05's out-of-bounds emulation depends on it). */
int oracle_debug_search_pb(int bx, int by, const float* px, const uint8_t* tp, int pc, int pb_in_batch,
                           or_symbolic* out) {
    or_desc* d = or_desc_get(bx, by);
    or_block b;
    if (!d || dbg_block(d, px, tp, pc, &b)) return -1;
    search_pb(d, &b, pc, pb_in_batch, out, NULL);
    return 0;
}

/* passes 001-003: k-means assignment of the block's texels to pc clusters. */
int oracle_debug_kmeans(int bx, int by, const float* px, int pc, uint8_t* assign) {
    or_desc* d = or_desc_get(bx, by);
    or_block b;
    if (!d || pc < 2 || pc > 4 || dbg_block(d, px, NULL, 1, &b)) return -1;
    kmeans(d, &b, pc, assign);
    return d->texel_count;
}

/* pass004: mismatch of an assignment against packed partitioning pidx. */
int oracle_debug_mismatch(int bx, int by, const uint8_t* assign, int pc, int pidx) {
    or_desc* d = or_desc_get(bx, by);
    if (!d || pc < 2 || pc > 4 || pidx < 0 || pidx >= OR_PARTITIONINGS) return -1;
    return (int)mismatch(d, assign, pc, pidx);
}

/* pass005: the first 128 entries of the partial bitonic ordering when only
 * entry 0 of the nsel selected partitionings has a non-zero score. */
int oracle_debug_ordering(int nsel, uint32_t score0, uint32_t* ord128) {
    ordering(nsel, score0, ord128);
    return 128;
}

/* pass006 for packed partitioning `entry` of count pc: {uncorrelated,
 * same-chroma} line-fit errors. */
int oracle_debug_eval_candidate(int bx, int by, const float* px, int pc, int entry, float* eu, float* es) {
    or_desc* d = or_desc_get(bx, by);
    or_block b;
    if (!d || pc < 2 || pc > 4 || entry < 0 || entry >= OR_PARTITIONINGS || dbg_block(d, px, NULL, 1, &b)) return -1;
    eval_candidate(d, &b, &d->part[(pc - 2) * OR_PARTITIONINGS + entry], pc, eu, es);
    return 0;
}

/* pass007: the 4 positions chosen from 128 {uncorrelated, same-chroma} errors. */
int oracle_debug_select_positions(const float* eu, const float* es, int* pos) {
    select_positions(eu, es, pos);
    return 0;
}
