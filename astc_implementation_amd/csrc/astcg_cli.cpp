// Command-line encoder with the reference CLI's arguments
// (code/main.cpp:225-343: `webgpu_astc <input> <output.astc> <block_x> <block_y>`).
//
// Reads uncompressed TGA (24/32-bit, either origin), binary PPM (P6) and
// PAM (P7, RGB or RGB_ALPHA) images; writes a .astc file with the 16-byte
// header of code/astc_store.cpp:5-61.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/astcg.h"

namespace {

bool read_file(const char* path, std::vector<uint8_t>& data) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    data.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

bool load_tga(const std::vector<uint8_t>& d, int& w, int& h, std::vector<uint8_t>& rgba) {
    if (d.size() < 18 || d[2] != 2) return false;  // uncompressed true-colour only
    const int idlen = d[0];
    w = d[12] | (d[13] << 8);
    h = d[14] | (d[15] << 8);
    const int bpp = d[16];
    const bool top_down = (d[17] & 0x20) != 0;
    if ((bpp != 24 && bpp != 32) || w <= 0 || h <= 0) return false;
    const size_t px = bpp / 8, off = 18 + idlen + (d[1] ? (size_t)(d[5] | (d[6] << 8)) * ((d[7] + 7) / 8) : 0);
    if (d.size() < off + (size_t)w * h * px) return false;
    rgba.resize((size_t)w * h * 4);
    for (int y = 0; y < h; ++y) {
        const int sy = top_down ? y : h - 1 - y;
        for (int x = 0; x < w; ++x) {
            const uint8_t* s = &d[off + ((size_t)sy * w + x) * px];
            uint8_t* o = &rgba[((size_t)y * w + x) * 4];
            o[0] = s[2]; o[1] = s[1]; o[2] = s[0]; o[3] = px == 4 ? s[3] : 255;
        }
    }
    return true;
}

// Parses the whitespace/comment separated header tokens of a PNM/PAM file.
size_t pnm_token(const std::vector<uint8_t>& d, size_t i, std::string& tok) {
    tok.clear();
    while (i < d.size()) {
        if (d[i] == '#') { while (i < d.size() && d[i] != '\n') ++i; continue; }
        if (isspace(d[i])) { ++i; continue; }
        break;
    }
    while (i < d.size() && !isspace(d[i])) tok += (char)d[i++];
    return i;
}

bool load_pnm(const std::vector<uint8_t>& d, int& w, int& h, std::vector<uint8_t>& rgba) {
    if (d.size() < 3 || d[0] != 'P') return false;
    std::string tok;
    size_t i = 2;
    int depth = 3, maxval = 255;
    if (d[1] == '6') {
        i = pnm_token(d, i, tok); w = atoi(tok.c_str());
        i = pnm_token(d, i, tok); h = atoi(tok.c_str());
        i = pnm_token(d, i, tok); maxval = atoi(tok.c_str());
        ++i;
    } else if (d[1] == '7') {
        for (;;) {
            i = pnm_token(d, i, tok);
            if (tok == "ENDHDR" || tok.empty()) break;
            std::string v;
            if (tok == "WIDTH") { i = pnm_token(d, i, v); w = atoi(v.c_str()); }
            else if (tok == "HEIGHT") { i = pnm_token(d, i, v); h = atoi(v.c_str()); }
            else if (tok == "DEPTH") { i = pnm_token(d, i, v); depth = atoi(v.c_str()); }
            else if (tok == "MAXVAL") { i = pnm_token(d, i, v); maxval = atoi(v.c_str()); }
            else if (tok == "TUPLTYPE") { i = pnm_token(d, i, v); }
        }
        ++i;
    } else {
        return false;
    }
    if (w <= 0 || h <= 0 || maxval != 255 || (depth != 3 && depth != 4)) return false;
    if (d.size() < i + (size_t)w * h * depth) return false;
    rgba.resize((size_t)w * h * 4);
    for (size_t p = 0; p < (size_t)w * h; ++p) {
        rgba[p * 4 + 0] = d[i + p * depth + 0];
        rgba[p * 4 + 1] = d[i + p * depth + 1];
        rgba[p * 4 + 2] = d[i + p * depth + 2];
        rgba[p * 4 + 3] = depth == 4 ? d[i + p * depth + 3] : 255;
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    // options after the reference CLI's four arguments (code/main.cpp:225-340):
    // --partitions N (1..4), --void-extent, --dual-plane
    astcg_params p;
    astcg_default_params(&p);
    bool ok = argc >= 5;
    for (int i = 5; ok && i < argc; ++i) {
        if (!std::strcmp(argv[i], "--void-extent")) p.flags |= ASTCG_FLAG_VOID_EXTENT;
        else if (!std::strcmp(argv[i], "--dual-plane")) p.flags |= ASTCG_FLAG_DUAL_PLANE;
        else if (!std::strcmp(argv[i], "--partitions") && i + 1 < argc) p.max_partitions = (uint32_t)std::atoi(argv[++i]);
        else ok = false;
    }
    if (!ok) {
        std::fprintf(stderr,
                     "usage: %s <input.tga|ppm|pam> <output.astc> <block_x> <block_y> "
                     "[--partitions N] [--void-extent] [--dual-plane]\n",
                     argv[0]);
        return 1;
    }
    const int bx = std::atoi(argv[3]), by = std::atoi(argv[4]);
    std::vector<uint8_t> file, rgba;
    int w = 0, h = 0;
    if (!read_file(argv[1], file)) { std::fprintf(stderr, "cannot read %s\n", argv[1]); return 1; }
    if (!load_tga(file, w, h, rgba) && !load_pnm(file, w, h, rgba)) {
        std::fprintf(stderr, "unsupported image format: %s\n", argv[1]);
        return 1;
    }
    astcg_encoder* enc = nullptr;
    if (astcg_create(0, &enc) != ASTCG_OK) { std::fprintf(stderr, "no HIP device\n"); return 1; }
    astcg_status st = astcg_configure(enc, (uint32_t)w, (uint32_t)h, (uint32_t)bx, (uint32_t)by, &p);
    if (st != ASTCG_OK) {
        std::fprintf(stderr, "configure: %s (%s)\n", astcg_status_string(st), astcg_last_error(enc));
        astcg_destroy(enc);
        return 1;
    }
    uint32_t nb = 0;
    astcg_block_counts(enc, nullptr, nullptr, &nb);
    std::vector<uint8_t> out((size_t)nb * 16);
    st = astcg_encode(enc, rgba.data(), rgba.size(), out.data(), out.size());
    if (st == ASTCG_OK) st = astcg_store_image((uint32_t)bx, (uint32_t)by, (uint32_t)w, (uint32_t)h, out.data(), out.size(), argv[2]);
    if (st != ASTCG_OK) std::fprintf(stderr, "encode: %s (%s)\n", astcg_status_string(st), astcg_last_error(enc));
    astcg_destroy(enc);
    return st == ASTCG_OK ? 0 : 1;
}
