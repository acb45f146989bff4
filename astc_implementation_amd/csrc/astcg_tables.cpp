// Spec-derived tables and per-footprint metadata for the gfx950 encoder.
//
// Reference behaviour followed (Jure-Z/ASTC_implementation):
//   block-mode decode        code/metadata_structures.cpp:72-173
//   decimation tables        code/metadata_structures.cpp:176-336
//   mode selection           code/compression_pipeline.cpp:146-198
//   k-means texel subset     code/metadata_structures.cpp:3-70
//   partition tables         code/partition_tables.cpp:16-436
//   angular tables           code/metadata_structures.cpp:463-481
//   lookup tables            shaders/pass07:12-37, pass11:12-63, pass14:29,
//                            pass16:13, code/physical_compression.cpp:44-622
#include "astcg_tables.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

namespace astcg {

static const int kLevels[21] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256};

void quant_btq(int q, int* bits, int* trits, int* quints) {
    int n = kLevels[q];
    *trits = (n % 3 == 0);
    *quints = (n % 5 == 0);
    int base = *trits ? n / 3 : (*quints ? n / 5 : n);
    int b = 0;
    while ((1 << b) < base) ++b;
    *bits = b;
}

int ise_bitcount(int count, int quant) {
    if (quant < 0 || quant > 20) return 1024;
    int b, t, u;
    quant_btq(quant, &b, &t, &u);
    return b * count + (t ? (8 * count + 4) / 5 : 0) + (u ? (7 * count + 2) / 3 : 0);
}

// Bit replication of a bits-wide value to `width` bits.
static int replicate_bits(int v, int bits, int width) {
    int out = 0;
    for (int filled = 0; filled < width;) {
        int take = std::min(bits, width - filled);
        out |= (v >> (bits - take)) << (width - filled - take);
        filled += take;
    }
    return out;
}

int weight_unquant(int q, int code) {
    int b, t, u;
    quant_btq(q, &b, &t, &u);
    int v;
    if (!t && !u) {
        v = replicate_bits(code, b, 6);
    } else if (b == 0) {
        static const int t3[3] = {0, 32, 63}, q5[5] = {0, 16, 32, 47, 63};
        v = t ? t3[code] : q5[code];
    } else {
        const int d = code >> b;
        const int bit_a = code & 1, bit_b = (code >> 1) & 1, bit_c = (code >> 2) & 1;
        const int A = bit_a ? 0x7F : 0;
        int B = 0, C;
        if (t) {
            C = b == 1 ? 50 : (b == 2 ? 23 : 11);
            if (b == 2) B = (bit_b << 6) | (bit_b << 2) | bit_b;
            else if (b == 3) B = (bit_c << 6) | (bit_b << 5) | (bit_c << 1) | bit_b;
        } else {
            C = b == 1 ? 28 : 13;
            if (b == 2) B = (bit_b << 6) | (bit_b << 1);
        }
        v = ((d * C + B) ^ A);
        v = (A & 0x20) | (v >> 2);
    }
    return v > 32 ? v + 1 : v;
}

int color_unquant(int q, int code) {
    int b, t, u;
    quant_btq(q, &b, &t, &u);
    if (!t && !u) return replicate_bits(code, b, 8);
    const int d = code >> b;
    int bit[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < b; ++i) bit[i] = (code >> i) & 1;
    const int A = bit[0] ? 0x1FF : 0;
    int B = 0, C = 0;
    // ASTC spec colour unquantisation constants (trits: C = 204/93/44/22/11/5,
    // quints: C = 113/54/26/13/6) with the B bit patterns of the spec table.
    const int bb = bit[1], c = bit[2], dd = bit[3], e = bit[4], f = bit[5];
    if (t) {
        switch (b) {
            case 1: C = 204; break;
            case 2: C = 93; B = (bb << 8) | (bb << 4) | (bb << 2) | (bb << 1); break;
            case 3: C = 44; B = (c << 8) | (bb << 7) | (c << 3) | (bb << 2) | (c << 1) | bb; break;
            case 4: C = 22; B = (dd << 8) | (c << 7) | (bb << 6) | (dd << 2) | (c << 1) | bb; break;
            case 5: C = 11; B = (e << 8) | (dd << 7) | (c << 6) | (bb << 5) | (e << 1) | dd; break;
            default: C = 5; B = (f << 8) | (e << 7) | (dd << 6) | (c << 5) | (bb << 4) | f; break;
        }
    } else {
        switch (b) {
            case 1: C = 113; break;
            case 2: C = 54; B = (bb << 8) | (bb << 3) | (bb << 2); break;
            case 3: C = 26; B = (c << 8) | (bb << 7) | (c << 2) | (bb << 1) | c; break;
            case 4: C = 13; B = (dd << 8) | (c << 7) | (bb << 6) | (dd << 1) | c; break;
            default: C = 6; B = (e << 8) | (dd << 7) | (c << 6) | (bb << 5) | e; break;
        }
    }
    int v = (d * C + B) ^ A;
    return (A & 0x80) | (v >> 2);
}

// Spec trit block decode: 8 packed bits -> 5 trits.
static void decode_trits(int T, int out[5]) {
    int C;
    if (((T >> 2) & 7) == 7) {
        C = (((T >> 5) & 7) << 2) | (T & 3);
        out[4] = 2;
        out[3] = 2;
    } else {
        C = T & 0x1F;
        if (((T >> 5) & 3) == 3) {
            out[4] = 2;
            out[3] = (T >> 7) & 1;
        } else {
            out[4] = (T >> 7) & 1;
            out[3] = (T >> 5) & 3;
        }
    }
    const int c0 = C & 1, c1 = (C >> 1) & 1, c2 = (C >> 2) & 1, c3 = (C >> 3) & 1, c4 = (C >> 4) & 1;
    if ((C & 3) == 3) {
        out[2] = 2;
        out[1] = c4;
        out[0] = (c3 << 1) | (c2 & (c3 ^ 1));
    } else if (((C >> 2) & 3) == 3) {
        out[2] = 2;
        out[1] = 2;
        out[0] = C & 3;
    } else {
        out[2] = c4;
        out[1] = (C >> 2) & 3;
        out[0] = (c1 << 1) | (c0 & (c1 ^ 1));
    }
}

// Spec quint block decode: 7 packed bits -> 3 quints.
static void decode_quints(int Q, int out[3]) {
    const int q0 = Q & 1, q3 = (Q >> 3) & 1, q4 = (Q >> 4) & 1;
    if (((Q >> 1) & 3) == 3 && ((Q >> 5) & 3) == 0) {
        out[2] = (q0 << 2) | ((q4 & (q0 ^ 1)) << 1) | (q3 & (q0 ^ 1));
        out[1] = 4;
        out[0] = 4;
        return;
    }
    int C;
    if (((Q >> 1) & 3) == 3) {
        out[2] = 4;
        C = (((Q >> 3) & 3) << 3) | ((~(Q >> 5) & 3) << 1) | q0;
    } else {
        out[2] = (Q >> 5) & 3;
        C = Q & 0x1F;
    }
    if ((C & 7) == 5) {
        out[1] = 4;
        out[0] = (C >> 3) & 3;
    } else {
        out[1] = (C >> 3) & 3;
        out[0] = C & 7;
    }
}

static QuantTables build_quant_tables() {
    QuantTables t;
    std::memset(&t, 0, sizeof(t));
    for (int q = 0; q < 12; ++q) {
        const int n = kLevels[q];
        std::vector<std::pair<int, int>> lv;
        for (int c = 0; c < n; ++c) lv.push_back({weight_unquant(q, c), c});
        std::sort(lv.begin(), lv.end());
        for (int r = 0; r < n; ++r) {
            t.quant_to_unquant[q][r] = (uint8_t)lv[r].first;
            t.weight_scramble[q][r] = (uint8_t)lv[r].second;
        }
        for (int r = 0; r < n; ++r) {
            const int prev = lv[std::max(r - 1, 0)].first, next = lv[std::min(r + 1, n - 1)].first;
            t.prev_next[q][lv[r].first] = (uint16_t)(prev | (next << 8));
        }
    }
    for (int q = 4; q < 21; ++q) {
        const int n = kLevels[q];
        std::vector<int> lv(n);
        for (int c = 0; c < n; ++c) lv[c] = color_unquant(q, c);
        for (int i = 0; i < 512; ++i) {
            int best = 0;
            for (int c = 1; c < n; ++c) {
                const int dc = std::abs(2 * lv[c] - i), db = std::abs(2 * lv[best] - i);
                if (dc < db || (dc == db && lv[c] < lv[best])) best = c;
            }
            t.color_u2q[q - 4][i] = (uint8_t)lv[best];
        }
        for (int v = 0; v < 256; ++v) {
            int best = 0;
            for (int c = 1; c < n; ++c) {
                const int dc = std::abs(lv[c] - v), db = std::abs(lv[best] - v);
                if (dc < db || (dc == db && c > best)) best = c;
            }
            t.color_pquant[q - 4][v] = (uint8_t)best;
        }
    }
    for (int row = 0; row < 10; ++row)
        for (int bits = 0; bits < 128; ++bits) {
            int best = -1;
            for (int q = 0; row > 0 && q < 21; ++q)
                if (ise_bitcount(2 * row, q) <= bits) best = q;
            t.quant_mode[row][bits] = (int8_t)best;
        }
    // Inverse of the spec decoders; for codes that decode to the same digits
    // the largest code is kept.
    for (int T = 0; T < 256; ++T) {
        int d[5];
        decode_trits(T, d);
        t.int_of_trits[(((d[4] * 3 + d[3]) * 3 + d[2]) * 3 + d[1]) * 3 + d[0]] = (uint8_t)T;
    }
    for (int Q = 0; Q < 128; ++Q) {
        int d[3];
        decode_quints(Q, d);
        t.int_of_quints[(d[2] * 5 + d[1]) * 5 + d[0]] = (uint8_t)Q;
    }
    for (int v = 0; v < 256; ++v) t.pixel_scale[v] = ((float)v / 255.0f) * 65536.0f;
    for (int T = 0; T < 256; ++T) {
        int d[5];
        decode_trits(T, d);
        t.trit_digits[T] = (uint16_t)(d[0] | (d[1] << 2) | (d[2] << 4) | (d[3] << 6) | (d[4] << 8));
    }
    for (int Q = 0; Q < 128; ++Q) {
        int d[3];
        decode_quints(Q, d);
        t.quint_digits[Q] = (uint16_t)(d[0] | (d[1] << 3) | (d[2] << 6));
    }
    for (int q = 4; q < 21; ++q)
        for (int c = 0; c < kLevels[q]; ++c) t.color_unq[q - 4][c] = (uint8_t)color_unquant(q, c);
    for (int q = 0; q < 12; ++q)
        for (int c = 0; c < kLevels[q]; ++c) t.weight_unq[q][c] = (uint8_t)weight_unquant(q, c);
    return t;
}

const QuantTables& quant_tables() {
    static const QuantTables tables = build_quant_tables();
    return tables;
}

// ---------------------------------------------------------------------------
// Footprint metadata

namespace {

struct ModeDesc {
    int xw, yw, dual, quant, wbits;
};

// ASTC 2D block-mode field decode (code/metadata_structures.cpp:72-173).
bool decode_mode(int mode, ModeDesc* m) {
    int base_q = (mode >> 4) & 1;
    int H = (mode >> 9) & 1, D = (mode >> 10) & 1;
    const int A = (mode >> 5) & 3;
    if (mode & 3) {
        base_q |= (mode & 3) << 1;
        int B = (mode >> 7) & 3;
        switch ((mode >> 2) & 3) {
            case 0: m->xw = B + 4; m->yw = A + 2; break;
            case 1: m->xw = B + 8; m->yw = A + 2; break;
            case 2: m->xw = A + 2; m->yw = B + 8; break;
            default:
                B &= 1;
                if (mode & 0x100) { m->xw = B + 2; m->yw = A + 2; }
                else { m->xw = A + 2; m->yw = B + 6; }
                break;
        }
    } else {
        base_q |= ((mode >> 2) & 3) << 1;
        if (((mode >> 2) & 3) == 0) return false;
        const int B = (mode >> 9) & 3;
        switch ((mode >> 7) & 3) {
            case 0: m->xw = 12; m->yw = A + 2; break;
            case 1: m->xw = A + 2; m->yw = 12; break;
            case 2: m->xw = A + 6; m->yw = B + 6; D = 0; H = 0; break;
            default:
                if (((mode >> 5) & 3) == 0) { m->xw = 6; m->yw = 10; }
                else if (((mode >> 5) & 3) == 1) { m->xw = 10; m->yw = 6; }
                else return false;
                break;
        }
    }
    const int wc = m->xw * m->yw * (D + 1);
    m->quant = (base_q - 2) + 6 * H;
    m->dual = D;
    m->wbits = ise_bitcount(wc, m->quant);
    return wc <= kMaxWeights && m->wbits >= 24 && m->wbits <= 96;
}

uint32_t partition_hash(uint32_t v) {
    v ^= v >> 15;
    v *= 0xEEDE0891u;
    v ^= v >> 5;
    v += v << 16;
    v ^= v >> 7;
    v ^= v >> 3;
    v ^= v << 6;
    v ^= v >> 17;
    return v;
}

// ASTC procedural partition function (code/partition_tables.cpp:120-241).
int partition_of(int seed, int x, int y, int pcount, bool small_block) {
    if (small_block) { x <<= 1; y <<= 1; }
    seed += (pcount - 1) * 1024;
    const uint32_t r = partition_hash((uint32_t)seed);
    uint8_t s[12];
    const int sh_src[12] = {0, 4, 8, 12, 16, 20, 24, 28, 18, 22, 26, 30};
    for (int i = 0; i < 11; ++i) s[i] = (r >> sh_src[i]) & 0xF;
    s[11] = ((r >> 30) | (r << 2)) & 0xF;
    for (int i = 0; i < 12; ++i) s[i] = (uint8_t)(s[i] * s[i]);
    int sh1, sh2;
    if (seed & 1) { sh1 = (seed & 2) ? 4 : 5; sh2 = pcount == 3 ? 6 : 5; }
    else { sh1 = pcount == 3 ? 6 : 5; sh2 = (seed & 2) ? 4 : 5; }
    const int sh3 = (seed & 0x10) ? sh1 : sh2;
    for (int i = 0; i < 8; ++i) s[i] >>= (i & 1) ? sh2 : sh1;
    for (int i = 8; i < 12; ++i) s[i] >>= sh3;
    // z = 0 for 2D blocks, so the z seeds drop out
    int a = (s[0] * x + s[1] * y + (int)(r >> 14)) & 0x3F;
    int b = (s[2] * x + s[3] * y + (int)(r >> 10)) & 0x3F;
    int c = (s[4] * x + s[5] * y + (int)(r >> 6)) & 0x3F;
    int d = (s[6] * x + s[7] * y + (int)(r >> 2)) & 0x3F;
    if (pcount <= 3) d = 0;
    if (pcount <= 2) c = 0;
    if (pcount <= 1) b = 0;
    if (a >= b && a >= c && a >= d) return 0;
    if (b >= c && b >= d) return 1;
    if (c >= d) return 2;
    return 3;
}

void build_decimation(Footprint& f, int dm, int xw, int yw, std::vector<std::vector<std::pair<int, int>>>& w2t) {
    const int xt = f.bx, yt = f.by, tc = f.texel_count;
    w2t.assign(xw * yw, {});
    for (int y = 0; y < yt; ++y)
        for (int x = 0; x < xt; ++x) {
            const int texel = y * xt + x;
            const int xwp = (((1024 + xt / 2) / (xt - 1)) * x * (xw - 1) + 32) >> 6;
            const int ywp = (((1024 + yt / 2) / (yt - 1)) * y * (yw - 1) + 32) >> 6;
            const int xf = xwp & 0xF, yf = ywp & 0xF, base = (xwp >> 4) + (ywp >> 4) * xw;
            const int idx[4] = {base, base + 1, base + xw, base + xw + 1};
            int w[4];
            w[3] = (xf * yf + 8) >> 4;
            w[1] = xf - w[3];
            w[2] = yf - w[3];
            w[0] = 16 - xf - yf + w[3];
            int n = 0;
            for (int i = 0; i < 4; ++i) {
                if (!w[i]) continue;
                const size_t slot = ((size_t)dm * tc + texel) * 4 + n;
                f.t2w_idx[slot] = (uint8_t)idx[i];
                f.t2w_w[slot] = (float)w[i] * (1.0f / 16.0f);
                ++n;
                w2t[idx[i]].push_back({texel, w[i]});
            }
            f.t2w_count[(size_t)dm * tc + texel] = (uint8_t)n;
        }
}

void build_partitions(Footprint& f) {
    const int tc = f.texel_count;
    const bool small_block = tc < 32;
    f.part_of_texel.assign((size_t)3 * kPartitionings * tc, 0);
    f.part_seed.assign(3 * kPartitionings, 0);
    f.part_counts.assign(3 * kPartitionings * 4, 0);
    f.coverage.assign(3 * kPartitionings * 4, 0);
    f.seed_to_packed.assign(3 * kPartitionings, 0xFFFF);
    std::vector<uint8_t> pot(tc);
    for (int pc = 2; pc <= 4; ++pc) {
        // Generate all seeds, then keep valid and canonically distinct patterns
        // first (the "selected" prefix), followed by all the others
        // (code/partition_tables.cpp:337-413).
        std::vector<std::vector<uint8_t>> pat(kPartitionings, std::vector<uint8_t>(tc));
        std::vector<int> counts(kPartitionings * 4, 0);
        std::vector<std::vector<uint8_t>> canon(kPartitionings, std::vector<uint8_t>(tc));
        std::vector<int> nparts(kPartitionings, 0);
        for (int s = 0; s < kPartitionings; ++s) {
            for (int y = 0; y < f.by; ++y)
                for (int x = 0; x < f.bx; ++x) {
                    const int p = partition_of(s, x, y, pc, small_block);
                    pat[s][y * f.bx + x] = (uint8_t)p;
                    counts[s * 4 + p]++;
                }
            int np = 0;
            while (np < 4 && counts[s * 4 + np] != 0) ++np;
            nparts[s] = np;
            int remap[4] = {-1, -1, -1, -1}, next = 0;
            for (int t = 0; t < tc; ++t) {
                int& m = remap[pat[s][t]];
                if (m < 0) m = next++;
                canon[s][t] = (uint8_t)m;
            }
        }
        std::vector<int> order;
        std::vector<char> built(kPartitionings, 0);
        for (int s = 0; s < kPartitionings; ++s) {
            if (nparts[s] != pc) continue;
            bool dup = false;
            for (int o : order)
                if (canon[o] == canon[s]) { dup = true; break; }
            if (dup) continue;
            order.push_back(s);
            built[s] = 1;
        }
        f.count_selected[pc - 1] = (int)order.size();
        for (int s = 0; s < kPartitionings; ++s)
            if (!built[s]) order.push_back(s);
        for (int k = 0; k < kPartitionings; ++k) {
            const int s = order[k];
            const size_t e = (size_t)(pc - 2) * kPartitionings + k;
            std::memcpy(&f.part_of_texel[e * tc], pat[s].data(), tc);
            f.part_seed[e] = (uint16_t)s;
            for (int i = 0; i < 4; ++i) f.part_counts[e * 4 + i] = (uint8_t)counts[s * 4 + i];
            f.seed_to_packed[(pc - 2) * kPartitionings + s] = (uint16_t)k;
            const int nk = std::min(tc, 64);
            for (int i = 0; i < nk; ++i)
                f.coverage[e * 4 + pat[s][f.kmeans_texels[i]]] |= 1ull << i;
        }
    }
}

// pass005_partition_ordering.wgsl:45-94: the network only compares pairs whose
// lower index is below 256; reproduce it for score[0] in {0, >0}, all other
// selected scores 0 and the rest ~0u.
void build_orderings(Footprint& f) {
    for (int pc = 2; pc <= 4; ++pc)
        for (int variant = 0; variant < 2; ++variant) {
            std::vector<uint32_t> score(1024), index(1024);
            for (int i = 0; i < 1024; ++i) {
                index[i] = i;
                score[i] = i < f.count_selected[pc - 1] ? (i == 0 ? (uint32_t)variant : 0u) : 0xFFFFFFFFu;
            }
            for (int k = 2; k <= 1024; k <<= 1)
                for (int j = k >> 1; j > 0; j >>= 1)
                    for (int l = 0; l < 256; ++l) {
                        const int p = l ^ j;
                        if (p <= l) continue;
                        const bool up = (l & k) == 0;
                        const bool swap = score[l] > score[p] || (score[l] == score[p] && index[l] > index[p]);
                        if (swap == up) {
                            std::swap(score[l], score[p]);
                            std::swap(index[l], index[p]);
                        }
                    }
            for (int i = 0; i < 128; ++i) f.ordering[pc - 2][variant][i] = (uint16_t)index[i];
        }
}

void build_kmeans_texels(Footprint& f) {
    const int tc = f.texel_count;
    if (tc <= 64) {
        for (int i = 0; i < tc; ++i) f.kmeans_texels[i] = (uint32_t)i;
        return;  // remaining entries stay 0
    }
    uint64_t s0 = 0xfaf9e171cea1ec6bULL, s1 = 0xf1b318cc06af5d71ULL;
    std::vector<char> seen(tc, 0);
    for (int n = 0; n < 64;) {
        const uint64_t res = s0 + s1;
        const uint64_t x = s1 ^ s0;
        s0 = ((s0 << 24) | (s0 >> 40)) ^ x ^ (x << 16);
        s1 = (x << 37) | (x >> 27);
        const int texel = (int)((uint8_t)res % tc);
        if (!seen[texel]) {
            seen[texel] = 1;
            f.kmeans_texels[n++] = (uint32_t)texel;
        }
    }
}

std::unique_ptr<Footprint> build_footprint(int bx, int by) {
    auto f = std::make_unique<Footprint>();
    f->bx = bx;
    f->by = by;
    const int tc = bx * by;
    f->texel_count = tc;
    // Enumerate block modes: packed decimation modes in order of first use.
    std::map<int, int> grid_to_dm;
    std::vector<std::pair<int, int>> dm_grid;
    std::vector<int> dm_refprec;
    struct Mode { int raw, dm, quant, wbits, dual; };
    std::vector<Mode> modes;
    f->raw_wc.assign(2048, 0);
    f->raw_quant.assign(2048, 0);
    f->raw_wbits.assign(2048, 0);
    for (int raw = 0; raw < 2048; ++raw) {
        ModeDesc m;
        if (!decode_mode(raw, &m) || m.xw > bx || m.yw > by) continue;
        if ((m.dual ? 109 : 111) - m.wbits <= 0) continue;
        const int key = m.yw * 16 + m.xw;
        auto it = grid_to_dm.find(key);
        int dm;
        if (it == grid_to_dm.end()) {
            dm = (int)dm_grid.size();
            grid_to_dm[key] = dm;
            dm_grid.push_back({m.xw, m.yw});
            dm_refprec.push_back(0);
        } else {
            dm = it->second;
        }
        if (!m.dual) dm_refprec[dm] |= 1 << m.quant;
        modes.push_back({raw, dm, m.quant, m.wbits, m.dual});
        f->raw_wc[raw] = (uint8_t)(m.xw * m.yw * (m.dual + 1));
        f->raw_quant[raw] = (uint8_t)m.quant;
        f->raw_wbits[raw] = (uint8_t)m.wbits;
    }
    // Trial decimation modes: those with a 1-plane mode at weight quant <= QUANT_12.
    std::vector<int> dm_trial(dm_grid.size(), -1);
    std::vector<int> trial_dm;
    for (size_t d = 0; d < dm_grid.size(); ++d)
        if (dm_refprec[d] & 0xFF) {
            dm_trial[d] = (int)trial_dm.size();
            trial_dm.push_back((int)d);
        }
    f->ndm = (int)trial_dm.size();
    f->dm.resize(f->ndm);
    f->t2w_count.assign((size_t)f->ndm * tc, 0);
    f->t2w_idx.assign((size_t)f->ndm * tc * 4, 0);
    f->t2w_w.assign((size_t)f->ndm * tc * 4, 0.0f);
    f->w2t_count.assign((size_t)f->ndm * 64, 0);
    f->w2t_off.assign((size_t)f->ndm * 64, 0);
    std::vector<std::vector<std::pair<int, int>>> w2t;
    for (int t = 0; t < f->ndm; ++t) {
        const int xw = dm_grid[trial_dm[t]].first, yw = dm_grid[trial_dm[t]].second, wc = xw * yw;
        build_decimation(*f, t, xw, yw, w2t);
        DmInfo& di = f->dm[t];
        di.texel_count = (uint8_t)tc;
        di.weight_count = (uint8_t)wc;
        di.wx = (uint8_t)xw;
        di.wy = (uint8_t)yw;
        int maxprec = -1;
        for (int q = 0; q < 12; ++q) {
            const int bits = ise_bitcount(wc, q);
            if (bits >= 24 && bits <= 96) maxprec = q;
        }
        static const int steps[12] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32};
        const int mp = std::max(0, std::min(maxprec, 7));
        di.max_quant_level = (uint8_t)mp;
        di.max_angular_steps = (uint8_t)steps[mp];
        di.max_quant_steps = (uint8_t)steps[mp];
        di.w2t_base = (uint32_t)f->w2t_idx.size();
        int rel = 0;
        for (int w = 0; w < wc; ++w) {
            f->w2t_count[t * 64 + w] = (uint8_t)w2t[w].size();
            f->w2t_off[t * 64 + w] = (uint16_t)rel;
            for (auto& e : w2t[w]) {
                f->w2t_idx.push_back((uint8_t)e.first);
                f->w2t_w.push_back((float)e.second * (1.0f / 16.0f));
                ++rel;
            }
        }
    }
    // Trial block modes: 1-plane modes whose decimation mode is a trial mode,
    // in raw-mode order.
    f->raw_to_bm.assign(2048, -1);
    for (const Mode& m : modes) {
        if (m.dual || dm_trial[m.dm] < 0) continue;
        BmInfo b{};
        b.raw_mode = (uint16_t)m.raw;
        b.dm = (uint8_t)dm_trial[m.dm];
        b.quant = (uint8_t)m.quant;
        b.weight_bits = (uint8_t)m.wbits;
        f->raw_to_bm[m.raw] = (int16_t)f->bm.size();
        f->bm.push_back(b);
    }
    f->nbm = (int)f->bm.size();
    // Dual-plane modes for the optional dual-plane search: a trial grid, a
    // legal weight bit count and room for at least QUANT_6 colour values
    // after the 17 mode / partition / CEM bits and the 2 plane-selector bits.
    for (const Mode& m : modes) {
        if (!m.dual || dm_trial[m.dm] < 0 || m.wbits < 24 || m.wbits > 96 || f->raw_wc[m.raw] > 64) continue;
        const int avail = 128 - 17 - 2 - m.wbits;
        int cq6 = -1, cq8 = -1;
        for (int q = 4; q <= 20; ++q) {
            if (ise_bitcount(6, q) <= avail) cq6 = q;
            if (ise_bitcount(8, q) <= avail) cq8 = q;
        }
        if (cq6 < 0) continue;
        DualMode d{};
        d.raw_mode = (uint16_t)m.raw;
        d.dm = (uint8_t)dm_trial[m.dm];
        d.quant = (uint8_t)m.quant;
        d.cq_rgb = (int8_t)cq6;
        d.cq_rgba = (int8_t)cq8;
        f->dual.push_back(d);
    }
    build_kmeans_texels(*f);
    build_partitions(*f);
    build_orderings(*f);
    const float kPi = 3.14159265358979323846f;
    f->sin_table.resize(1024 * 16);
    f->cos_table.resize(1024 * 16);
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 1024; ++j) {
            const float angle = (2.0f * kPi / (1024 - 1.0f)) * (float)(i + 1) * (float)j;
            f->sin_table[j * 16 + i] = sinf(angle);
            f->cos_table[j * 16 + i] = cosf(angle);
        }
    return f;
}

}  // namespace

bool valid_footprint(int bx, int by) {
    static const int fp[14][2] = {{4, 4}, {5, 4}, {5, 5}, {6, 5}, {6, 6}, {8, 5}, {8, 6},
                                  {10, 5}, {10, 6}, {8, 8}, {10, 8}, {10, 10}, {12, 10}, {12, 12}};
    for (auto& p : fp)
        if (p[0] == bx && p[1] == by) return true;
    return false;
}

const Footprint* footprint(int bx, int by) {
    if (!valid_footprint(bx, by)) return nullptr;
    static std::mutex lock;
    static std::map<int, std::unique_ptr<Footprint>> cache;
    std::lock_guard<std::mutex> g(lock);
    auto& slot = cache[bx * 16 + by];
    if (!slot) slot = build_footprint(bx, by);
    return slot.get();
}

}  // namespace astcg
