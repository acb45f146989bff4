// Exercises the C oracle and the product's host codec and tables under
// AddressSanitizer and UndefinedBehaviorSanitizer (built and run by
// tests/test_sanitizers.py with gcc / g++ -fsanitize=address,undefined).
// Returns 0 when every check holds; the sanitizers abort on a finding.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "astc_oracle.h"
#include "astcg_codec.h"
#include "astcg_tables.h"

namespace {

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint32_t rnd() {
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return (uint32_t)(g_state >> 32);
}

int fails = 0;
void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "check failed: %s\n", what);
        ++fails;
    }
}

// gradient + noise + a hard edge, alpha varying when `alpha`
std::vector<uint8_t> image(int w, int h, bool alpha) {
    std::vector<uint8_t> px((size_t)w * h * 4);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            uint8_t* p = &px[((size_t)y * w + x) * 4];
            p[0] = (uint8_t)(x * 255 / (w - 1));
            p[1] = (uint8_t)((y * 200 / (h - 1) + (rnd() & 15)) & 0xFF);
            p[2] = x + y < (w + h) / 2 ? 30 : 220;
            p[3] = alpha ? (uint8_t)(128 + (int)(rnd() & 63) - 32) : 255;
        }
    return px;
}

}  // namespace

int main() {
    const int fps[14][2] = {{4, 4}, {5, 4}, {5, 5}, {6, 5}, {6, 6}, {8, 5}, {8, 6},
                            {8, 8}, {10, 5}, {10, 6}, {10, 8}, {10, 10}, {12, 10}, {12, 12}};
    // the product's footprint tables
    for (const auto& f : fps) {
        const astcg::Footprint* fp = astcg::footprint(f[0], f[1]);
        check(fp && fp->texel_count == f[0] * f[1] && fp->nbm > 0, "footprint tables");
    }
    // the oracle's encoder (every policy) at four footprints, blocks decoded
    // by the product's host decoder
    const int enc_fps[4][2] = {{4, 4}, {6, 6}, {8, 5}, {12, 12}};
    for (const auto& f : enc_fps)
        for (int policy : {0, 2, 3}) {
            const int w = 3 * f[0] + 1, h = 2 * f[1] + 3;
            const std::vector<uint8_t> img = image(w, h, policy == 2);
            const int nb = ((w + f[0] - 1) / f[0]) * ((h + f[1] - 1) / f[1]);
            std::vector<uint8_t> blocks((size_t)nb * 16);
            std::vector<or_symbolic> sym(nb);
            const int32_t seeds[3][6] = {{1, 2, 3, 4, 5, 6}, {7, 8, 9, 10, 11, 12}, {13, 14, 15, 16, 17, 18}};
            or_params prm;
            std::memset(&prm, 0, sizeof(prm));
            prm.max_partitions = 4;
            prm.threads = 1;
            prm.policy = policy;
            prm.search_limit = policy == 2 ? 50 : 0;
            prm.seed_list = &seeds[0][0];
            prm.seed_list_len = 6;
            prm.mismatch_candidates = 8;
            check(oracle_encode(img.data(), w, h, f[0], f[1], &prm, blocks.data(), sym.data(), nullptr) == 0,
                  "oracle_encode");
            std::vector<uint8_t> tex((size_t)f[0] * f[1] * 4);
            for (int b = 0; b < nb; ++b) {
                check(astcg::decode_block_rgba8(f[0], f[1], &blocks[(size_t)b * 16], tex.data()) == 0,
                      "decode of an encoded block");
                astcg_symbolic_block sb;
                check(astcg::physical_to_symbolic_host(f[0], f[1], &blocks[(size_t)b * 16], &sb) == 0,
                      "physical_to_symbolic of an encoded block");
                uint8_t again[16];
                check(astcg::symbolic_to_physical_host(f[0], f[1], sb, again) == 0 &&
                          std::memcmp(again, &blocks[(size_t)b * 16], 16) == 0,
                      "symbolic round trip");
            }
        }
    // random 128-bit blocks through the decoder and the symbolic reader
    // (every error path), at every footprint
    for (const auto& f : fps) {
        std::vector<uint8_t> tex((size_t)f[0] * f[1] * 4);
        for (int i = 0; i < 400; ++i) {
            uint8_t blk[16];
            for (int j = 0; j < 16; ++j) blk[j] = (uint8_t)rnd();
            (void)astcg::decode_block_rgba8(f[0], f[1], blk, tex.data());
            astcg_symbolic_block sb;
            (void)astcg::physical_to_symbolic_host(f[0], f[1], blk, &sb);
        }
    }
    // ISE round trips at every quant level and length
    for (int q = 0; q < 21; ++q)
        for (int n = 1; n <= 64; n += 7) {
            uint8_t in[64], out[64], packed[40];
            const int levels[21] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256};
            for (int j = 0; j < n; ++j) in[j] = (uint8_t)(rnd() % levels[q]);
            std::memset(packed, 0, sizeof(packed));
            const int bits = astcg::ise_bitcount(n, q);
            if (bits > 8 * 32) continue;
            astcg::encode_ise_host(q, n, in, packed, 0);
            astcg::decode_ise_host(q, n, packed, 0, out);
            check(std::memcmp(in, out, n) == 0, "ISE round trip");
            uint8_t ref[40];
            std::memset(ref, 0, sizeof(ref));
            oracle_encode_ise(q, n, in, ref, 0);
            check(std::memcmp(ref, packed, (bits + 7) / 8) == 0, "ISE equals the oracle's");
        }
    std::printf("sanitizer driver: %d failed checks\n", fails);
    return fails ? 1 : 0;
}
