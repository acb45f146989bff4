// C ABI of the gfx950 ASTC encoder (declared in include/astcg.h).
//
// The host side mirrors ASTCEncoder (code/compression_pipeline.cpp:200-535):
// configure() builds the per-footprint metadata and uploads it once; encode()
// runs the search on the device in chunks of blocks and packs the 128-bit
// blocks on the device.  Unlike the reference there is no per-batch host
// round trip: the only device->host copy is the finished block stream.
// Footprint tables are cached per encoder, so batches of textures with
// several footprints reuse them; the working buffers grow on demand.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <cstring>
#include <fstream>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/astcg.h"
#include "../../include/astcg_predictor.h"
#include "astcg_codec.h"
#include "astcg_device.h"
#include "astcg_tables.h"

using namespace astcg;

// Working memory of astcg_encode_device_noalloc: a fixed device array (each
// GPU the library runs on holds its own copy until the process ends).  Host
// code never passes `scratch` itself to a kernel or a copy: its name is only
// a host-side stand-in; the device address comes from hipGetSymbolAddress,
// taken once per device in astcg_create and kept in g_scratch.
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device
}  // namespace

// Device tables of one footprint.
struct FpState {
    const Footprint* fp = nullptr;
    void* d_tables = nullptr;
    DevTables T{};
    int wstride = 0;  // LDS row of quantised weights: max trial weight count, multiple of 4
};

// Candidate predictor of one footprint (LEARNED policy).
struct PredState {
    int nin = 0, hidden = 0, nout = 0, nmodes = 0;
    void* w16[3] = {nullptr, nullptr, nullptr};  // bf16 W1, W2, W3
    float* bias[3] = {nullptr, nullptr, nullptr};
    int* d_head = nullptr;  // [15]: head offsets, lengths, per-head top-k counts
    void* w3_seed = nullptr;  // bf16 [3][1024][hidden]: seed-head columns of W3, transposed
    // fused inference (astcg_mlp_infer; hidden width 256): transposed bf16
    // weights, the output layer reduced to the count and mode heads
    void* fw[3] = {nullptr, nullptr, nullptr};
    float* fb3 = nullptr;
    astcg_learned_params lp{};
};

void free_pred(PredState& P) {
    for (int i = 0; i < 3; ++i) {
        if (P.w16[i]) (void)hipFree(P.w16[i]);
        if (P.bias[i]) (void)hipFree(P.bias[i]);
        P.w16[i] = nullptr;
        P.bias[i] = nullptr;
    }
    if (P.d_head) (void)hipFree(P.d_head);
    if (P.w3_seed) (void)hipFree(P.w3_seed);
    P.d_head = nullptr;
    P.w3_seed = nullptr;
    for (int i = 0; i < 3; ++i) {
        if (P.fw[i]) (void)hipFree(P.fw[i]);
        P.fw[i] = nullptr;
    }
    if (P.fb3) (void)hipFree(P.fb3);
    P.fb3 = nullptr;
}

struct astcg_encoder {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    std::map<int, FpState> fps;  // keyed by bx * 16 + by
    // working buffers, grown on demand
    void* d_chunk = nullptr;
    size_t chunk_cap = 0;
    ChunkBuffers B{};
    const FpState* layout_fp = nullptr;  // footprint and chunk size B is laid out for
    int layout_chunk = 0;
    SymBest* d_best = nullptr;
    size_t best_cap = 0;
    // single-image configuration
    bool configured = false;
    uint32_t width = 0, height = 0, bx = 0, by = 0, blocks_x = 0, blocks_y = 0, nblocks = 0;
    astcg_params params{};
    const FpState* cur = nullptr;
    int chunk_blocks = 0;
    int chunk_limit = 0;  // automatic chunk size cap after an allocation failure (0 = none)
    uint8_t* d_rgba = nullptr;  // host-buffer staging
    uint8_t* d_out = nullptr;
    size_t rgba_cap = 0, out_cap = 0;
    // batch texture tables
    TexDesc* d_tex = nullptr;
    size_t tex_cap = 0;
    std::vector<TexDesc> h_tex;
    // pinned staging of the texture table: the upload is enqueued without a
    // host wait; the next batch waits only for that copy (tex_copied) before
    // it rewrites the staging buffer
    TexDesc* h_tex_pinned = nullptr;
    size_t tex_pinned_cap = 0;
    hipEvent_t tex_copied = nullptr;
    bool tex_copy_pending = false;
    std::map<int, PredState> preds;  // keyed like fps
    void* d_pwork = nullptr;         // predictor inference workspace
    size_t pwork_cap = 0;
    void* d_plan = nullptr;          // search plan buffers
    size_t plan_cap = 0;
    long long fb_total = 0;                 // blocks searched under the LEARNED policy since the last reset
    unsigned long long* d_fb = nullptr;     // device count of their fallback blocks
    void* d_pcseed = nullptr;  // label buffers (LabelBufs) with ASTCG_FLAG_LABELS
    size_t pcseed_cap = 0;
    struct EventProfiler* prof = nullptr;
    void* scratch = nullptr;  // this device's static scratch (astcg_encode_device_noalloc)
    // ASTCG_FLAG_GRAPH: instantiated launch chains, keyed by everything the
    // chain's launches depend on (most recently used last)
    std::vector<std::pair<std::vector<uint8_t>, hipGraphExec_t>> graphs;
};

constexpr size_t kMaxGraphs = 16;

// Records a start/stop event pair around every kernel launch; the pairs are
// resolved (and the pool recycled) when the times are queried.
struct EventProfiler : Profiler {
    struct Rec { int kernel; hipEvent_t a, b; };
    std::vector<hipEvent_t> pool;
    std::vector<Rec> recs;
    double total_ms[K_COUNT] = {};
    long long launches[K_COUNT] = {};
    hipEvent_t open = nullptr;
    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    void begin(int, hipStream_t s) override {
        open = take();
        (void)hipEventRecord(open, s);
    }
    void end(int kernel, hipStream_t s) override {
        hipEvent_t b = take();
        (void)hipEventRecord(b, s);
        recs.push_back({kernel, open, b});
    }
    void resolve() {
        for (Rec& r : recs) {
            float ms = 0.0f;
            (void)hipEventSynchronize(r.b);
            (void)hipEventElapsedTime(&ms, r.a, r.b);
            total_ms[r.kernel] += ms;
            launches[r.kernel] += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        recs.clear();
    }
    ~EventProfiler() override {
        resolve();
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

namespace {

astcg_status fail(astcg_encoder* e, astcg_status s, const std::string& msg) {
    if (e) e->last_error = msg;
    return s;
}

astcg_status hip_fail(astcg_encoder* e, hipError_t err, const char* what) {
    (void)hipGetLastError();  // reported through the status; later calls (ours or the caller's) start clean
    return fail(e, err == hipErrorOutOfMemory ? ASTCG_ERR_OUT_OF_MEMORY : ASTCG_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(err));
}

void release_all(astcg_encoder* e) {
    for (auto& kv : e->fps)
        if (kv.second.d_tables) (void)hipFree(kv.second.d_tables);
    e->fps.clear();
    if (e->d_chunk) (void)hipFree(e->d_chunk);
    if (e->d_best) (void)hipFree(e->d_best);
    if (e->d_rgba) (void)hipFree(e->d_rgba);
    if (e->d_out) (void)hipFree(e->d_out);
    if (e->d_tex) (void)hipFree(e->d_tex);
    if (e->tex_copied) (void)hipEventSynchronize(e->tex_copied);
    if (e->h_tex_pinned) (void)hipHostFree(e->h_tex_pinned);
    if (e->tex_copied) (void)hipEventDestroy(e->tex_copied);
    e->h_tex_pinned = nullptr;
    e->tex_pinned_cap = 0;
    e->tex_copied = nullptr;
    e->tex_copy_pending = false;
    if (e->d_pcseed) (void)hipFree(e->d_pcseed);
    for (auto& kv : e->preds) free_pred(kv.second);
    e->preds.clear();
    for (auto& g : e->graphs) (void)hipGraphExecDestroy(g.second);
    e->graphs.clear();
    if (e->d_fb) (void)hipFree(e->d_fb);
    e->d_fb = nullptr;
    if (e->d_pwork) (void)hipFree(e->d_pwork);
    if (e->d_plan) (void)hipFree(e->d_plan);
    e->d_pwork = e->d_plan = nullptr;
    e->pwork_cap = e->plan_cap = 0;
    e->d_pcseed = nullptr;
    e->pcseed_cap = 0;
    e->d_chunk = nullptr;
    e->d_best = nullptr;
    e->d_rgba = e->d_out = nullptr;
    e->d_tex = nullptr;
    e->chunk_cap = e->best_cap = e->rgba_cap = e->out_cap = e->tex_cap = 0;
    e->layout_fp = nullptr;
    e->cur = nullptr;
    delete e->prof;
    e->prof = nullptr;
    e->configured = false;
}

// Grows a device buffer to at least `bytes` (contents are not kept).  Work
// already queued may still use the old buffer, so the device is drained
// before it is freed.
hipError_t grow(void** p, size_t* cap, size_t bytes) {
    if (*cap >= bytes && *p) return hipSuccess;
    if (*p) {
        hipError_t err = hipDeviceSynchronize();
        if (err != hipSuccess) return err;
        (void)hipFree(*p);
    }
    *p = nullptr;
    *cap = 0;
    hipError_t err = hipMalloc(p, std::max<size_t>(bytes, 256));
    if (err == hipSuccess) *cap = std::max<size_t>(bytes, 256);
    else (void)hipGetLastError();  // reported through the return value, not left for later calls
    return err;
}

// Packs all footprint tables into one device allocation.
struct Layout {
    size_t size = 0;
    template <typename T>
    size_t add(size_t count) {
        size = (size + 255) & ~size_t(255);
        const size_t off = size;
        size += count * sizeof(T);
        return off;
    }
};

hipError_t upload_tables(FpState& F) {
    const Footprint& f = *F.fp;
    const int tc = f.texel_count;
    // CSR of block modes grouped by decimation mode
    std::vector<uint16_t> off(f.ndm + 1, 0), list;
    for (int m = 0; m < f.ndm; ++m) {
        off[m] = (uint16_t)list.size();
        for (int k = 0; k < f.nbm; ++k)
            if (f.bm[k].dm == m) list.push_back((uint16_t)k);
    }
    off[f.ndm] = (uint16_t)list.size();
    // quantised weights of a trial mode: a region of npb rows, one row of
    // the mode's weight count rounded up to 4 per partitioned block, so the
    // rows are written and read as whole dwords
    std::vector<uint16_t> qw_off(std::max(f.nbm, 1), 0);
    int qw_total = 0;
    for (int k = 0; k < f.nbm; ++k) {
        qw_off[k] = (uint16_t)qw_total;
        qw_total += (f.dm[f.bm[k].dm].weight_count + 3) & ~3;
    }
    std::vector<uint16_t> ordering(3 * 2 * 128);
    for (int p = 0; p < 3; ++p)
        for (int v = 0; v < 2; ++v)
            for (int i = 0; i < 128; ++i) ordering[(p * 2 + v) * 128 + i] = f.ordering[p][v][i];
    Layout L;
    const size_t o_dm = L.add<DmInfo>(f.dm.size());
    const size_t o_bm = L.add<BmInfo>(f.bm.size());
    const size_t o_off = L.add<uint16_t>(off.size());
    const size_t o_list = L.add<uint16_t>(std::max<size_t>(list.size(), 1));
    const size_t o_t2wc = L.add<uint8_t>(f.t2w_count.size());
    const size_t o_t2wi = L.add<uint8_t>(f.t2w_idx.size());
    const size_t o_t2ww = L.add<float>(f.t2w_w.size());
    const size_t o_w2tc = L.add<uint8_t>(f.w2t_count.size());
    const size_t o_w2to = L.add<uint16_t>(f.w2t_off.size());
    const size_t o_w2ti = L.add<uint8_t>(std::max<size_t>(f.w2t_idx.size(), 1));
    const size_t o_w2tw = L.add<float>(std::max<size_t>(f.w2t_w.size(), 1));
    const size_t o_pot = L.add<uint8_t>(f.part_of_texel.size());
    const size_t o_seed = L.add<uint16_t>(f.part_seed.size());
    const size_t o_cnt = L.add<uint8_t>(f.part_counts.size());
    const size_t o_cov = L.add<uint64_t>(f.coverage.size());
    const size_t o_ord = L.add<uint16_t>(ordering.size());
    const size_t o_km = L.add<uint32_t>(64);
    const size_t o_sin = L.add<float>(f.sin_table.size());
    const size_t o_cos = L.add<float>(f.cos_table.size());
    const size_t o_r2b = L.add<int16_t>(f.raw_to_bm.size());
    const size_t o_rwc = L.add<uint8_t>(f.raw_wc.size());
    const size_t o_rq = L.add<uint8_t>(f.raw_quant.size());
    const size_t o_qwo = L.add<uint16_t>(qw_off.size());
    const size_t o_s2p = L.add<uint16_t>(f.seed_to_packed.size());
    const size_t o_dual = L.add<DualMode>(std::max<size_t>(f.dual.size(), 1));
    // decimation modes grouped by weight count for k_dm_bm's launches
    std::vector<uint8_t> dm_order;
    int group_start[kDmGroups + 1] = {}, group_w[kDmGroups] = {};
    {
        const int limits[kDmGroups] = {16, 24, 32, 1 << 30};
        for (int g = 0; g < kDmGroups; ++g) {
            group_w[g] = 4;
            group_start[g] = (int)dm_order.size();
            for (int d = 0; d < f.ndm; ++d) {
                const int wc = f.dm[d].weight_count;
                if (wc <= limits[g] && (g == 0 || wc > limits[g - 1])) {
                    dm_order.push_back((uint8_t)d);
                    group_w[g] = std::max(group_w[g], (wc + 3) & ~3);
                }
            }
        }
        group_start[kDmGroups] = (int)dm_order.size();
    }
    const size_t o_dmo = L.add<uint8_t>(std::max<size_t>(dm_order.size(), 1));
    std::vector<uint8_t> host(L.size, 0);
    auto put = [&](size_t o, const void* src, size_t bytes) { if (bytes) std::memcpy(host.data() + o, src, bytes); };
    put(o_dm, f.dm.data(), f.dm.size() * sizeof(DmInfo));
    put(o_bm, f.bm.data(), f.bm.size() * sizeof(BmInfo));
    put(o_off, off.data(), off.size() * 2);
    put(o_list, list.data(), list.size() * 2);
    put(o_t2wc, f.t2w_count.data(), f.t2w_count.size());
    put(o_t2wi, f.t2w_idx.data(), f.t2w_idx.size());
    put(o_t2ww, f.t2w_w.data(), f.t2w_w.size() * 4);
    put(o_w2tc, f.w2t_count.data(), f.w2t_count.size());
    put(o_w2to, f.w2t_off.data(), f.w2t_off.size() * 2);
    put(o_w2ti, f.w2t_idx.data(), f.w2t_idx.size());
    put(o_w2tw, f.w2t_w.data(), f.w2t_w.size() * 4);
    put(o_pot, f.part_of_texel.data(), f.part_of_texel.size());
    put(o_seed, f.part_seed.data(), f.part_seed.size() * 2);
    put(o_cnt, f.part_counts.data(), f.part_counts.size());
    put(o_cov, f.coverage.data(), f.coverage.size() * 8);
    put(o_ord, ordering.data(), ordering.size() * 2);
    put(o_km, f.kmeans_texels, 64 * 4);
    put(o_sin, f.sin_table.data(), f.sin_table.size() * 4);
    put(o_cos, f.cos_table.data(), f.cos_table.size() * 4);
    put(o_r2b, f.raw_to_bm.data(), f.raw_to_bm.size() * 2);
    put(o_rwc, f.raw_wc.data(), f.raw_wc.size());
    put(o_rq, f.raw_quant.data(), f.raw_quant.size());
    put(o_qwo, qw_off.data(), qw_off.size() * 2);
    put(o_s2p, f.seed_to_packed.data(), f.seed_to_packed.size() * 2);
    put(o_dual, f.dual.data(), f.dual.size() * sizeof(DualMode));
    put(o_dmo, dm_order.data(), dm_order.size());
    hipError_t err = hipMalloc(&F.d_tables, L.size);
    if (err != hipSuccess) return err;
    err = hipMemcpy(F.d_tables, host.data(), L.size, hipMemcpyHostToDevice);
    if (err != hipSuccess) return err;
    uint8_t* b = static_cast<uint8_t*>(F.d_tables);
    DevTables& T = F.T;
    T.dm = reinterpret_cast<const DmInfo*>(b + o_dm);
    T.bm = reinterpret_cast<const BmInfo*>(b + o_bm);
    T.dm_bm_off = reinterpret_cast<const uint16_t*>(b + o_off);
    T.dm_bm_list = reinterpret_cast<const uint16_t*>(b + o_list);
    T.t2w_count = b + o_t2wc;
    T.t2w_idx = b + o_t2wi;
    T.t2w_w = reinterpret_cast<const float*>(b + o_t2ww);
    T.w2t_count = b + o_w2tc;
    T.w2t_off = reinterpret_cast<const uint16_t*>(b + o_w2to);
    T.w2t_idx = b + o_w2ti;
    T.w2t_w = reinterpret_cast<const float*>(b + o_w2tw);
    T.part_of_texel = b + o_pot;
    T.part_seed = reinterpret_cast<const uint16_t*>(b + o_seed);
    T.part_counts = b + o_cnt;
    T.coverage = reinterpret_cast<const uint64_t*>(b + o_cov);
    T.ordering = reinterpret_cast<const uint16_t*>(b + o_ord);
    T.kmeans_texels = reinterpret_cast<const uint32_t*>(b + o_km);
    T.sin_t = reinterpret_cast<const float*>(b + o_sin);
    T.cos_t = reinterpret_cast<const float*>(b + o_cos);
    T.raw_to_bm = reinterpret_cast<const int16_t*>(b + o_r2b);
    T.raw_wc = b + o_rwc;
    T.raw_quant = b + o_rq;
    T.qw_off = reinterpret_cast<const uint16_t*>(b + o_qwo);
    T.seed_to_packed = reinterpret_cast<const uint16_t*>(b + o_s2p);
    T.dual = reinterpret_cast<const DualMode*>(b + o_dual);
    T.dm_order = b + o_dmo;
    for (int g = 0; g <= kDmGroups; ++g) T.dm_group_start[g] = group_start[g];
    for (int g = 0; g < kDmGroups; ++g) T.dm_group_w[g] = group_w[g];
    T.ndual = (int)f.dual.size();
    T.qw_stride = (qw_total + 15) & ~15;
    T.bx = f.bx;
    T.by = f.by;
    T.tc = tc;
    T.ndm = f.ndm;
    T.nbm = f.nbm;
    for (int i = 0; i < 4; ++i) T.count_selected[i] = f.count_selected[i];
    T.refine_stage_max_tc = kRefineStageMaxTc;
    T.ex_f4_max_tc = kExF4MaxTc;
    // from 36 texels per block a wave per listed block keeps up with a lane
    // per block on long lists too (6x6 learned 4096^2: -1 %); at 4x4 the
    // lane kernel wins once its waves fill the chip (+2 % without the cut)
    T.exw_max_list = f.texel_count >= kExwAllListsTc ? kExwAllLists : kExwMaxList;
    T.refine_queue = kRefineQueue;
    // grouping by format class pays where the refinement stages its texels
    // (4x4 4096^2 images: -10 % refinement time); unstaged, the classes split
    // neighbouring blocks' candidates apart and their texel reads lose cache
    // locality (6x6: +4 %)
    T.refine_classes = f.texel_count <= kRefineStageMaxTc ? 1 : 0;
    int maxw = 4;
    for (const DmInfo& d : f.dm) maxw = std::max<int>(maxw, d.weight_count);
    F.wstride = (maxw + 3) & ~3;
    return hipSuccess;
}

size_t chunk_bytes_per_block(const Footprint& f, size_t* offs, int nb) {
    Layout L;
    const size_t npb = (size_t)nb * 4;
    offs[0] = L.add<int16_t>(npb);
    offs[1] = L.add<Ideal>(npb);
    offs[2] = L.add<float>(npb * f.texel_count);
    offs[3] = L.add<float>(npb * kColorTabRows);
    offs[10] = L.add<uint16_t>(npb * kColorTabRows);
    offs[4] = L.add<Ccr>(npb * f.nbm);
    int qw_total = 0;
    for (int k = 0; k < f.nbm; ++k) qw_total += (f.dm[f.bm[k].dm].weight_count + 3) & ~3;
    offs[5] = L.add<uint8_t>(npb * (size_t)((qw_total + 15) & ~15));
    offs[6] = L.add<int16_t>(npb * 8);
    offs[7] = L.add<CandOut>(npb * 8);
    const size_t bins = (size_t)f.nbm * kRefineClasses;
    offs[8] = L.add<int>(2 * bins + 2);
    offs[9] = L.add<int>(npb * 8 + bins * 64);
    offs[11] = L.add<float>(npb * 8);
    offs[12] = L.add<uint32_t>((size_t)nb * f.texel_count);
    return L.size;
}

// Tables of footprint bx x by, uploaded on first use.
// ---- launch-shape tuning and its cache file ----------------------------------
// One line per GPU and footprint:
//   astcg-tune 2 <gcnArchName> <CU count> <bx>x<by> <refine_stage_max_tc> <ex_f4_max_tc> <exw_max_list>
//                <refine_queue> <refine_classes>
// (lines of format 1, without the last two fields, keep their defaults)
std::string tune_device_id(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown 0";
    std::string arch = prop.gcnArchName;
    for (char& c : arch)
        if (c == ' ') c = '_';
    return arch + " " + std::to_string(prop.multiProcessorCount);
}

bool tune_cache_lookup(const char* path, const std::string& dev, uint32_t bx, uint32_t by, astcg_tuning* t) {
    FILE* f = path ? std::fopen(path, "r") : nullptr;
    if (!f) return false;
    char line[512];
    bool found = false;
    const std::string want = dev + " " + std::to_string(bx) + "x" + std::to_string(by) + " ";
    while (std::fgets(line, sizeof(line), f)) {
        const std::string l(line);
        int version = 0;
        for (int v = 1; v <= 2; ++v) {
            const std::string head = "astcg-tune " + std::to_string(v) + " ";
            if (l.compare(0, head.size(), head) == 0 && l.compare(head.size(), want.size(), want) == 0) version = v;
        }
        if (!version) continue;
        unsigned r = 0, x = 0, w = 0, q = t->refine_queue, c = t->refine_classes;
        const int n = std::sscanf(l.c_str() + 13 + want.size(), "%u %u %u %u %u", &r, &x, &w, &q, &c);
        if (n == (version == 1 ? 3 : 5) && x <= 64 && q <= 64 && c <= 1) {
            t->refine_stage_max_tc = r;
            t->ex_f4_max_tc = x;
            t->exw_max_list = w;
            t->refine_queue = q;
            t->refine_classes = c;
            found = true;  // the last entry of the file wins
        }
    }
    std::fclose(f);
    return found;
}

void apply_tuning(FpState& F, const astcg_tuning& t) {
    F.T.refine_stage_max_tc = (int)std::min<uint32_t>(t.refine_stage_max_tc, 1u << 20);
    F.T.ex_f4_max_tc = (int)std::min<uint32_t>(t.ex_f4_max_tc, 64u);
    F.T.exw_max_list = (int)std::min<uint32_t>(t.exw_max_list, 0x7FFFFFFFu);
    F.T.refine_queue = (int)std::min<uint32_t>(t.refine_queue, 64u);
    F.T.refine_classes = t.refine_classes ? 1 : 0;
}

astcg_tuning tuning_of(const FpState& F) {
    astcg_tuning t;
    t.refine_stage_max_tc = (uint32_t)F.T.refine_stage_max_tc;
    t.ex_f4_max_tc = (uint32_t)F.T.ex_f4_max_tc;
    t.exw_max_list = (uint32_t)F.T.exw_max_list;
    t.refine_queue = (uint32_t)F.T.refine_queue;
    t.refine_classes = (uint32_t)F.T.refine_classes;
    return t;
}

astcg_status get_footprint_mut(astcg_encoder* e, uint32_t bx, uint32_t by, FpState** out) {
    const int key = (int)(bx * 16 + by);
    auto it = e->fps.find(key);
    if (it == e->fps.end()) {
        FpState F;
        F.fp = footprint((int)bx, (int)by);
        hipError_t err = upload_tables(F);
        if (err != hipSuccess) {
            if (F.d_tables) (void)hipFree(F.d_tables);
            return hip_fail(e, err, "table upload");
        }
        // launch shapes tuned earlier for this GPU, if the cache names them
        astcg_tuning t;
        const char* cache = std::getenv("ASTCG_TUNE_CACHE");
        if (cache && tune_cache_lookup(cache, tune_device_id(e->device), bx, by, &t)) apply_tuning(F, t);
        it = e->fps.emplace(key, F).first;
    }
    *out = &it->second;
    return ASTCG_OK;
}

astcg_status get_footprint(astcg_encoder* e, uint32_t bx, uint32_t by, const FpState** out) {
    FpState* F = nullptr;
    astcg_status st = get_footprint_mut(e, bx, by, &F);
    *out = F;
    return st;
}

// Blocks per chunk: keeps the per-chunk working set near 20 GiB, so a 4096^2
// image takes one chunk at every footprint (14.0 GiB at 4x4, less at larger
// footprints; up to 17 GiB at 4x4 with the learned policy's plan buffers for
// 64 predicted seeds).  Every partition count of a chunk ends with a few long
// refinement lanes that leave most of the GPU idle, so fewer, larger chunks
// pay that tail fewer times.  A requested size is capped so that the element
// indices of the per-PB arrays ([colour-table row or trial mode][PB]) stay
// below 2^31.
size_t plan_bytes(const Footprint& f, const PredState* P, int nb);

int chunk_size(const Footprint& f, uint32_t nblocks, uint32_t requested, const PredState* pred) {
    size_t offs[13];
    const size_t fixed = chunk_bytes_per_block(f, offs, 0) + plan_bytes(f, pred, 0);
    const size_t per_block = (chunk_bytes_per_block(f, offs, 1 << 16) + plan_bytes(f, pred, 1 << 16) - fixed) >> 16;
    const size_t max_chunk = (((size_t)1 << 31) - 1) / (4 * (size_t)std::max(kColorTabRows, f.nbm) + 8 * 4);
    size_t chunk = std::max<size_t>(1024, ((size_t(20) << 30) - fixed) / per_block);
    if (requested) chunk = requested;
    chunk = std::min<size_t>(chunk, max_chunk);
    return (int)std::max<size_t>(1, std::min<size_t>(chunk, nblocks));
}

// The chunk buffers of footprint F at `base`, with the offsets of
// chunk_bytes_per_block.
ChunkBuffers chunk_buffers_at(uint8_t* b, const FpState& F, const size_t* offs) {
    ChunkBuffers B;
    B.pb_part = reinterpret_cast<int16_t*>(b + offs[0]);
    B.ideal = reinterpret_cast<Ideal*>(b + offs[1]);
    B.wts = reinterpret_cast<float*>(b + offs[2]);
    B.ct_err = reinterpret_cast<float*>(b + offs[3]);
    B.ct_fmt = reinterpret_cast<uint16_t*>(b + offs[10]);
    B.ccr = reinterpret_cast<Ccr*>(b + offs[4]);
    B.qw = b + offs[5];
    B.top = reinterpret_cast<int16_t*>(b + offs[6]);
    B.cand = reinterpret_cast<CandOut*>(b + offs[7]);
    B.cand_err = reinterpret_cast<float*>(b + offs[11]);
    B.bin_count = reinterpret_cast<int*>(b + offs[8]);
    B.bin_cur = B.bin_count + (size_t)F.fp->nbm * kRefineClasses;
    B.work_total = B.bin_cur + (size_t)F.fp->nbm * kRefineClasses;
    B.tile_next = B.work_total + 1;
    B.work = reinterpret_cast<int*>(b + offs[9]);
    B.texels = reinterpret_cast<uint32_t*>(b + offs[12]);
    B.wstride = F.wstride;
    return B;
}

// Lays the chunk buffers out for footprint F and `chunk` blocks, growing the
// allocation if needed; the per-mode counters are cleared whenever the
// layout changes (k_bins clears them after every use otherwise).
astcg_status layout_chunk(astcg_encoder* e, const FpState& F, int chunk, hipStream_t s) {
    if (e->layout_fp == &F && e->layout_chunk == chunk) return ASTCG_OK;
    size_t offs[13];
    const size_t total = chunk_bytes_per_block(*F.fp, offs, chunk);
    hipError_t err = grow(&e->d_chunk, &e->chunk_cap, total);
    if (err != hipSuccess) {
        e->layout_fp = nullptr;  // the old buffers are gone
        e->layout_chunk = 0;
        return hip_fail(e, err, "chunk buffers");
    }
    ChunkBuffers& B = e->B;
    B = chunk_buffers_at(static_cast<uint8_t*>(e->d_chunk), F, offs);
    err = hipMemsetAsync(B.bin_count, 0, sizeof(int) * (2 * (size_t)F.fp->nbm * kRefineClasses + 2), s);
    if (err != hipSuccess) return hip_fail(e, err, "chunk buffers");
    e->layout_fp = &F;
    e->layout_chunk = chunk;
    return ASTCG_OK;
}

// Label buffers of ASTCG_FLAG_LABELS over nb blocks, in one allocation:
// candidate errors (float [3][4][nb]), best seeds, candidate seeds (int16).
size_t label_bytes(size_t nb) { return nb * (12 * sizeof(float) + 15 * sizeof(int16_t)); }

LabelBufs label_bufs(void* base, size_t nb) {
    LabelBufs l;
    l.cand_err = static_cast<float*>(base);
    l.best_seed = reinterpret_cast<int16_t*>(l.cand_err + 12 * nb);
    l.cand_seed = l.best_seed + 3 * nb;
    return l;
}

astcg_status plan_chunk(astcg_encoder* e, const FpState& F, const PredState& P, const ImageArgs& A,
                        const std::vector<TexDesc>* texs, hipStream_t s, SearchPlan* plan, bool reserve_only = false);

// Lays the chunk buffers out for `nblocks` blocks of footprint F (and, with a
// predictor, sizes the LEARNED policy's plan buffers and inference workspace
// for a chunk) and returns the chunk size used.  The automatic size
// (requested == 0) halves after an allocation failure (another process may
// hold device memory) down to 4096 blocks, and the encoder keeps the smaller
// size until the next configure or batched call; results do not depend on
// the chunk size.  A requested size is used as given.
astcg_status layout_fitting(astcg_encoder* e, const FpState& F, uint32_t nblocks, uint32_t requested, hipStream_t s,
                            int* chunk_out, const PredState* pred = nullptr) {
    int chunk = chunk_size(*F.fp, nblocks, requested, pred);
    if (!requested && e->chunk_limit > 0) chunk = std::min(chunk, e->chunk_limit);
    for (;;) {
        astcg_status st = layout_chunk(e, F, chunk, s);
        if (st == ASTCG_OK && pred) {
            ImageArgs A1{};
            A1.nb = chunk;
            SearchPlan unused;
            st = plan_chunk(e, F, *pred, A1, nullptr, s, &unused, true);
        }
        if (st == ASTCG_OK) {
            *chunk_out = chunk;
            return ASTCG_OK;
        }
        if (st != ASTCG_ERR_OUT_OF_MEMORY || requested || chunk <= 4096) return st;
        chunk = std::max(4096, chunk / 2);
        e->chunk_limit = chunk;
    }
}

constexpr int kInferBlocks = 32768;  // predictor inference sub-chunk
typedef const uint8_t* bfloat16_ptr;    // bf16 weights addressed bytewise

// Plan buffers and inference workspace of the LEARNED policy for a chunk of
// nb blocks: offsets into the plan allocation (o_*) and the workspace (w_*).
struct PlanLayout {
    int kmax = 4, words = 1, sub = 0;
    bool rank = false, prune_refine = false, fused = false, seed_topk = false;
    size_t o_idx = 0, o_prob = 0, o_mask = 0, o_slots = 0, o_fb = 0, o_list = 0, o_inv = 0, o_cnt = 0, o_h2all = 0,
           o_est = 0, plan_size = 0;
    size_t w_x = 0, w_z = 0, w_h = 0, w_z2 = 0, w_h2 = 0, w_l = 0, work_size = 0;
};

PlanLayout plan_layout(const Footprint& f, const PredState& P, int nb) {
    const astcg_learned_params& lp = P.lp;
    PlanLayout Y;
    Y.kmax = (int)std::max<uint32_t>(std::max<uint32_t>(lp.seeds, lp.modes), 4);
    Y.words = (f.nbm + 31) / 32;
    struct Lay { size_t size = 0; size_t add(size_t b) { size = (size + 255) & ~size_t(255); size_t o = size; size += b; return o; } } L;
    Y.o_idx = L.add(sizeof(int) * (size_t)nb * 5 * Y.kmax);
    Y.o_prob = L.add(sizeof(float) * (size_t)nb * 5 * Y.kmax);
    Y.o_mask = L.add(sizeof(uint32_t) * (size_t)nb * Y.words);
    Y.o_slots = L.add((size_t)nb);
    Y.o_fb = L.add((size_t)nb);
    Y.o_list = L.add(sizeof(int) * (size_t)nb * 4);
    Y.o_inv = L.add(sizeof(int) * (size_t)nb * 4);
    Y.o_cnt = L.add(sizeof(int) * 5);  // 3 list counts, the two need-list counts
    // ranking the compat positions, and the seed heads' top-k of the fused
    // path, read every block's second hidden layer
    Y.rank = P.w3_seed != nullptr;
    Y.o_h2all = Y.rank ? L.add(2 * (size_t)nb * P.hidden) : 0;
    Y.prune_refine = lp.refine_partitionings >= 1 && lp.refine_partitionings < 4;
    Y.o_est = Y.prune_refine ? L.add(sizeof(float) * (size_t)nb * 4) : 0;
    Y.plan_size = L.size;
    // inference workspace for one sub-chunk.  The fused kernel needs only the
    // features (2 * nin bytes per block) and runs over the whole chunk in one
    // launch, so its workgroups overlap their serial layer-by-layer latency;
    // the separate GEMM / GELU / top-k launches keep fp32 intermediates of the
    // 3-head output layer (about 16 KB per block) and run in sub-chunks.
    Y.fused = P.fw[0] && (lp.seeds == 0 || P.w3_seed);
    Y.seed_topk = Y.fused && lp.seeds > 0;
    const int H = P.hidden, NO = P.nout;
    Y.sub = Y.fused ? nb : std::min(nb, kInferBlocks);
    const int usub = Y.fused ? 0 : Y.sub;  // rows of the unfused intermediates
    Lay W;
    Y.w_x = W.add(2 * (size_t)Y.sub * P.nin);
    Y.w_z = W.add(4 * (size_t)usub * H);
    Y.w_h = W.add(2 * (size_t)usub * H);
    Y.w_z2 = W.add(4 * (size_t)usub * H);
    Y.w_h2 = W.add(2 * (size_t)usub * H);
    Y.w_l = W.add(4 * (size_t)usub * NO);
    Y.work_size = W.size;
    return Y;
}

// LEARNED policy: runs the predictor over the chunk's blocks and lays out the
// search plan (top-k of every head per block, then the plan kernels).  With
// reserve_only the buffers are only sized for A.nb blocks.
size_t plan_bytes(const Footprint& f, const PredState* P, int nb) {
    if (!P) return 0;
    const PlanLayout Y = plan_layout(f, *P, nb);
    return Y.plan_size + Y.work_size;
}

astcg_status plan_chunk(astcg_encoder* e, const FpState& F, const PredState& P, const ImageArgs& A,
                        const std::vector<TexDesc>* texs, hipStream_t s, SearchPlan* plan, bool reserve_only) {
    const astcg_learned_params& lp = P.lp;
    const int nb = A.nb;
    const PlanLayout Y = plan_layout(*F.fp, P, nb);
    const int kmax = Y.kmax, words = Y.words;
    hipError_t err = grow(&e->d_plan, &e->plan_cap, Y.plan_size);
    if (err != hipSuccess) return hip_fail(e, err, "plan buffers");
    err = grow(&e->d_pwork, &e->pwork_cap, Y.work_size);
    if (err != hipSuccess) return hip_fail(e, err, "predictor workspace");
    if (reserve_only) return ASTCG_OK;  // buffers sized, nothing enqueued
    uint8_t* pb = static_cast<uint8_t*>(e->d_plan);
    const size_t o_idx = Y.o_idx, o_prob = Y.o_prob, o_mask = Y.o_mask, o_slots = Y.o_slots, o_fb = Y.o_fb,
                 o_list = Y.o_list, o_inv = Y.o_inv, o_cnt = Y.o_cnt, o_h2all = Y.o_h2all, o_est = Y.o_est;
    const size_t o_x = Y.w_x, o_z = Y.w_z, o_h = Y.w_h, o_z2 = Y.w_z2, o_h2 = Y.w_h2, o_l = Y.w_l;
    const bool rank = Y.rank, prune_refine = Y.prune_refine, fused = Y.fused;
    const int H = P.hidden, NO = P.nout, sub = Y.sub;
    uint8_t* wb = static_cast<uint8_t*>(e->d_pwork);
    void* X = wb + o_x;
    float* Z1 = reinterpret_cast<float*>(wb + o_z);
    void* H1 = wb + o_h;
    float* Z2 = reinterpret_cast<float*>(wb + o_z2);
    void* H2 = wb + o_h2;
    float* LG = reinterpret_cast<float*>(wb + o_l);
    int* tidx = reinterpret_cast<int*>(pb + o_idx);
    float* tprob = reinterpret_cast<float*>(pb + o_prob);
    for (int b0 = 0; b0 < nb; b0 += sub) {
        const int n = std::min(sub, nb - b0);
        const int g0 = A.blk0 + b0;
        // features: one call per texture overlapping the blocks g0 .. g0 + n - 1
        for (const TexDesc& t : *texs) {
            const int lo = std::max(g0, t.first_block), hi = std::min(g0 + n, t.first_block + t.num_blocks);
            if (lo >= hi) continue;
            astcg_status st = astcg_mlp_features(t.rgba, t.width, t.height, (int)F.fp->bx, (int)F.fp->by,
                                                 lo - t.first_block, hi - lo,
                                                 static_cast<uint8_t*>(X) + 2 * (size_t)(lo - g0) * P.nin, P.nin, s);
            if (st != ASTCG_OK) return fail(e, st, "predictor features");
        }
        if (fused) {
            // one fused kernel: hidden layers, count and mode heads, top-k
            astcg_status st = astcg_mlp_infer(n, X, P.nin, P.nin, P.fw[0], P.bias[0], P.fw[1], P.bias[1], P.fw[2], P.fb3,
                                              4 + P.nmodes, P.nmodes, kmax, 4, (int)std::min<uint32_t>(lp.modes, P.nmodes),
                                              tidx + (size_t)b0 * 5 * kmax, tprob + (size_t)b0 * 5 * kmax,
                                              rank ? pb + o_h2all + 2 * (size_t)b0 * H : nullptr, s);
            if (st != ASTCG_OK) return fail(e, st, "predictor inference");
            continue;
        }
        astcg_status st = astcg_mlp_gemm(n, H, P.nin, X, P.nin, 0, P.w16[0], H, 0, Z1, H, P.bias[0], 0, s);
        if (st == ASTCG_OK) st = astcg_mlp_gelu(Z1, H1, (long long)n * H, s);
        if (st == ASTCG_OK) st = astcg_mlp_gemm(n, H, H, H1, H, 0, P.w16[1], H, 0, Z2, H, P.bias[1], 0, s);
        if (rank) H2 = pb + o_h2all + 2 * (size_t)b0 * H;
        if (st == ASTCG_OK) st = astcg_mlp_gelu(Z2, H2, (long long)n * H, s);
        // output layer: only the heads the plan reads (the seed heads are 3072
        // of the columns and unused with the compat positions)
        const int mode_off = 4 + 3 * 1024;
        const bfloat16_ptr w3 = static_cast<const uint8_t*>(P.w16[2]);
        if (lp.seeds == 0) {
            if (st == ASTCG_OK) st = astcg_mlp_gemm(n, 4, H, H2, H, 0, w3, NO, 0, LG, NO, P.bias[2], 0, s);
            if (st == ASTCG_OK)
                st = astcg_mlp_gemm(n, P.nmodes, H, H2, H, 0, w3 + 2 * (size_t)mode_off, NO, 0, LG + mode_off, NO,
                                    P.bias[2] + mode_off, 0, s);
        } else if (st == ASTCG_OK) {
            st = astcg_mlp_gemm(n, NO, H, H2, H, 0, w3, NO, 0, LG, NO, P.bias[2], 0, s);
        }
        if (st == ASTCG_OK)
            st = astcg_mlp_topk_heads(LG, n, NO, 5, P.d_head, P.d_head + 5, kmax, P.d_head + 10,
                                      tidx + (size_t)b0 * 5 * kmax, tprob + (size_t)b0 * 5 * kmax, s);
        if (st != ASTCG_OK) return fail(e, st, "predictor inference");
    }
    SearchPlan& pl = *plan;
    pl = SearchPlan{};
    pl.mode_mask = reinterpret_cast<uint32_t*>(pb + o_mask);
    pl.mask_words = words;
    pl.slots = pb + o_slots;
    pl.pb_list = reinterpret_cast<int*>(pb + o_list);
    pl.pb_inv = reinterpret_cast<int*>(pb + o_inv);
    pl.pb_count = reinterpret_cast<int*>(pb + o_cnt);
    pl.topk_idx = tidx;
    pl.topk_prob = tprob;
    pl.kmax = kmax;
    pl.seeds = (int)lp.seeds;
    pl.modes = (int)lp.modes;
    pl.nslots = (int)lp.slots;
    pl.fallback_below = lp.fallback_below;
    pl.count_below = lp.count_below;
    pl.fallback = pb + o_fb;
    pl.fb_count = e->d_fb;
    pl.keep_pb = (lp.seeds == 0 && lp.partitionings >= 1 && lp.partitionings < 4) ? (int)lp.partitionings : 4;
    pl.seed_topk = Y.seed_topk ? 1 : 0;
    pl.hidden = H;
    pl.h2 = rank ? reinterpret_cast<const uint16_t*>(pb + o_h2all) : nullptr;
    pl.w3_seed = static_cast<const uint16_t*>(P.w3_seed);
    pl.b3_seed = P.bias[2] + 4;
    pl.keep_refine = prune_refine ? (int)lp.refine_partitionings : 4;
    {
        // the reference's error limit (code/compression_pipeline.cpp:276-282, 330)
        const float texels = (float)(F.fp->bx * F.fp->by);
        const float db = 80.0f - 19.0f * (logf(texels) / logf(10.0f));
        const float weights_sum = 0.30f * 2.25f + 0.59f * 2.25f + 0.11f * 2.25f + 1.0f;  // astc.h:60-63
        const float limit = powf(0.1f, db * 0.1f) * 65535.0f * 65535.0f * texels * weights_sum;
        pl.error_limit = lp.early_out > 0.0f ? lp.early_out * limit : 0.0f;
        pl.full_limit = lp.full_search_above > 0.0f ? lp.full_search_above * limit : 0.0f;
        pl.count_limit = lp.count_above > 0.0f ? lp.count_above * limit : 0.0f;
        pl.hard_keeps_modes = lp.full_search_modes == 1 ? 1 : 0;
        pl.mis_candidates = lp.seeds > 4 ? (int)lp.mismatch_candidates : 0;
        pl.line_limit = lp.fallback_above > 0.0f ? lp.fallback_above * limit : 0.0f;
    }
    pl.best = e->d_best;
    pl.iterations = lp.rounds ? (int)lp.rounds : 6;
    pl.part_slots = lp.partitioned_slots ? (int)lp.partitioned_slots : (int)lp.slots;
    pl.pb_est = prune_refine ? reinterpret_cast<float*>(pb + o_est) : nullptr;
    return ASTCG_OK;
}

// Runs the whole search + pack over the blocks described by A (A.num_blocks
// blocks of one footprint), chunk by chunk, on stream s.  `pred` selects the
// LEARNED policy (texs: the textures of A, for the predictor's inputs).
astcg_status run_encode(astcg_encoder* e, const FpState& F, ImageArgs A, int max_pc, uint32_t chunk_req,
                        uint32_t flags, hipStream_t s, const LabelBufs* labels = nullptr, const PredState* pred = nullptr,
                        const std::vector<TexDesc>* texs = nullptr, int exhaustive = 0) {
    // the learned policy's candidates: 1 = every selected partitioning, N >= 2
    // = the first N in k-means mismatch order (partition_search_limit)
    if (pred && pred->lp.candidates >= 1) exhaustive = pred->lp.candidates == 1 ? kPartitionings : (int)pred->lp.candidates;
    int chunk = 0;
    astcg_status st = layout_fitting(e, F, (uint32_t)A.num_blocks, chunk_req, s, &chunk, pred);
    if (st != ASTCG_OK) return st;
    hipError_t err = grow(reinterpret_cast<void**>(&e->d_best), &e->best_cap, sizeof(SymBest) * (size_t)A.num_blocks);
    if (err != hipSuccess) return hip_fail(e, err, "symbolic block buffer");
    std::vector<TexDesc> one;
    if (pred && !texs) {
        one.push_back(A.t0);
        texs = &one;
    }
    if (pred) e->fb_total += A.num_blocks;
    // graph replay: capture needs a stream of the caller's (not the null
    // stream) and no per-launch profiling events
    const bool graph = (flags & ASTCG_FLAG_GRAPH) && s != nullptr && !(flags & ASTCG_FLAG_PROFILE);
    std::vector<uint8_t> key;
    if (graph) {
        // layout_fitting has sized every buffer (plan buffers included): no
        // allocation happens while capturing
        auto put = [&key](const void* p, size_t n) {
            const uint8_t* b = static_cast<const uint8_t*>(p);
            key.insert(key.end(), b, b + n);
        };
        const void* ptrs[] = {&F, pred, e->d_chunk, e->d_best, e->d_plan, e->d_pwork, e->d_fb,
                              labels ? labels->best_seed : nullptr, labels ? labels->cand_seed : nullptr,
                              labels ? labels->cand_err : nullptr};
        put(ptrs, sizeof(ptrs));
        const int ints[] = {max_pc, chunk, (int)flags, e->layout_chunk, exhaustive, F.T.refine_stage_max_tc,
                            F.T.ex_f4_max_tc, F.T.exw_max_list, F.T.refine_queue, F.T.refine_classes};
        put(ints, sizeof(ints));
        put(&A, sizeof(A));
        if (pred) put(&pred->lp, sizeof(pred->lp));
        if (texs) put(texs->data(), sizeof(TexDesc) * texs->size());
        for (size_t i = 0; i < e->graphs.size(); ++i) {
            if (e->graphs[i].first != key) continue;
            auto hit = e->graphs[i];
            e->graphs.erase(e->graphs.begin() + (long)i);
            e->graphs.push_back(hit);
            err = hipGraphLaunch(hit.second, s);
            if (err != hipSuccess) return hip_fail(e, err, "graph launch");
            return ASTCG_OK;
        }
        err = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
        if (err != hipSuccess) return hip_fail(e, err, "graph capture");
    }
    // the launch chain: every chunk's plan and search, then the packer
    st = ASTCG_OK;
    for (int blk0 = 0; blk0 < A.num_blocks && st == ASTCG_OK; blk0 += chunk) {
        A.blk0 = blk0;
        A.nb = std::min(chunk, A.num_blocks - blk0);
        SearchPlan plan;
        if (pred) st = plan_chunk(e, F, *pred, A, texs, s, &plan);
        if (st != ASTCG_OK) break;
        err = launch_chunk(F.T, A, e->B, e->d_best, max_pc, s, e->prof, labels, pred ? &plan : nullptr,
                           (flags & ASTCG_FLAG_DUAL_PLANE) != 0, exhaustive);
        if (err != hipSuccess) st = hip_fail(e, err, "search kernels");
    }
    if (st == ASTCG_OK) {
        err = launch_pack(F.T, A, e->d_best, s, e->prof, (flags & ASTCG_FLAG_VOID_EXTENT) != 0);
        if (err != hipSuccess) st = hip_fail(e, err, "pack kernel");
    }
    if (!graph) return st;
    hipGraph_t g = nullptr;
    err = hipStreamEndCapture(s, &g);
    if (st != ASTCG_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (err != hipSuccess) return hip_fail(e, err, "graph capture");
    hipGraphExec_t x = nullptr;
    err = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess) return hip_fail(e, err, "graph instantiate");
    if (e->graphs.size() >= kMaxGraphs) {
        (void)hipStreamSynchronize(s);
        (void)hipGraphExecDestroy(e->graphs.front().second);
        e->graphs.erase(e->graphs.begin());
    }
    e->graphs.emplace_back(std::move(key), x);
    err = hipGraphLaunch(x, s);
    if (err != hipSuccess) return hip_fail(e, err, "graph launch");
    return ASTCG_OK;
}

astcg_status check_params(astcg_encoder* e, const astcg_params* params, astcg_params* p) {
    astcg_default_params(p);
    if (params) *p = *params;
    if (p->max_partitions < 1 || p->max_partitions > 4)
        return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "max_partitions must be 1..4");
    if (p->policy != ASTCG_POLICY_COMPAT && p->policy != ASTCG_POLICY_LEARNED && p->policy != ASTCG_POLICY_EXHAUSTIVE)
        return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "unknown partition policy");
    if ((p->flags & ASTCG_FLAG_PROFILE) && !e->prof) e->prof = new EventProfiler();
    return ASTCG_OK;
}

}  // namespace

extern "C" {

void astcg_default_params(astcg_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->max_partitions = 4;
    p->policy = ASTCG_POLICY_COMPAT;
    p->chunk_blocks = 0;
}

astcg_status astcg_create(int hip_device, astcg_encoder** out) {
    if (!out) return ASTCG_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    hipError_t err = hipGetDeviceCount(&n);
    if (err != hipSuccess || hip_device < 0 || hip_device >= n) return ASTCG_ERR_DEVICE;
    auto* e = new astcg_encoder();
    e->device = hip_device;
    err = hipSetDevice(hip_device);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (err != hipSuccess) {
        delete e;
        return ASTCG_ERR_DEVICE;
    }
    upload_quant_tables(e->stream);
    err = hipStreamSynchronize(e->stream);
    if (err == hipSuccess && hip_device < kMaxDevices) {
        // the device address of this device's static scratch, taken once
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch[hip_device]) {
            void* p = nullptr;
            err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (err == hipSuccess) g_scratch[hip_device] = p;
        }
        e->scratch = g_scratch[hip_device];
    }
    if (err != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamDestroy(e->stream);
        delete e;
        return ASTCG_ERR_DEVICE;
    }
    *out = e;
    return ASTCG_OK;
}

astcg_status astcg_destroy(astcg_encoder* e) {
    if (!e) return ASTCG_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    release_all(e);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return ASTCG_OK;
}

astcg_status astcg_configure(astcg_encoder* e, uint32_t width, uint32_t height, uint32_t bx, uint32_t by,
                             const astcg_params* params) {
    if (!e) return ASTCG_ERR_INVALID_ARGUMENT;
    if (width == 0 || height == 0 || width >= (1u << 24) || height >= (1u << 24))
        return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "image dimensions must be in 1..2^24-1");
    if (!valid_footprint((int)bx, (int)by)) return fail(e, ASTCG_ERR_INVALID_FOOTPRINT, "not a 2D ASTC footprint");
    astcg_params p;
    astcg_status st = check_params(e, params, &p);
    if (st != ASTCG_OK) return st;
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    e->configured = false;
    const uint32_t bxn = (width + bx - 1) / bx, byn = (height + by - 1) / by;
    const uint64_t nb = (uint64_t)bxn * byn;
    if (nb > 0x7FFFFFFFull) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "too many blocks");
    const FpState* F = nullptr;
    st = get_footprint(e, bx, by, &F);
    if (st != ASTCG_OK) return st;
    if (p.policy == ASTCG_POLICY_LEARNED && !e->preds.count((int)(bx * 16 + by)))
        return fail(e, ASTCG_ERR_NOT_CONFIGURED, "LEARNED policy without a predictor for this footprint");
    e->width = width;
    e->height = height;
    e->bx = bx;
    e->by = by;
    e->blocks_x = bxn;
    e->blocks_y = byn;
    e->nblocks = (uint32_t)nb;
    e->params = p;
    e->cur = F;
    // allocate the working set now (plan buffers included under the LEARNED
    // policy) so that encode_device allocates nothing; a chunk cap left by an
    // earlier memory shortage is dropped, the full automatic size is tried
    // first
    e->chunk_limit = 0;
    const PredState* pred = nullptr;
    if (p.policy == ASTCG_POLICY_LEARNED) pred = &e->preds.find((int)(bx * 16 + by))->second;
    st = layout_fitting(e, *F, e->nblocks, p.chunk_blocks, e->stream, &e->chunk_blocks, pred);
    if (st != ASTCG_OK) return st;
    err = grow(reinterpret_cast<void**>(&e->d_best), &e->best_cap, sizeof(SymBest) * (size_t)e->nblocks);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) return hip_fail(e, err, "symbolic block buffer");
    e->configured = true;
    return ASTCG_OK;
}

astcg_status astcg_encode_device(astcg_encoder* e, const uint8_t* d_rgba, uint8_t* d_out, void* stream) {
    if (!e || !d_rgba || !d_out) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "null pointer");
    if (!e->configured) return fail(e, ASTCG_ERR_NOT_CONFIGURED, "encode before configure");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    ImageArgs A{};
    A.t0.rgba = d_rgba;
    A.t0.out = d_out;
    A.t0.width = (int)e->width;
    A.t0.height = (int)e->height;
    A.t0.blocks_x = (int)e->blocks_x;
    A.t0.first_block = 0;
    A.t0.num_blocks = (int)e->nblocks;
    A.ntex = 1;
    A.num_blocks = (int)e->nblocks;
    LabelBufs lab{};
    const LabelBufs* labels = nullptr;
    if (e->params.flags & ASTCG_FLAG_LABELS) {
        const size_t bytes = label_bytes(e->nblocks);
        err = grow(&e->d_pcseed, &e->pcseed_cap, bytes);
        if (err == hipSuccess) err = hipMemsetAsync(e->d_pcseed, 0xFF, bytes, s);
        if (err != hipSuccess) return hip_fail(e, err, "label buffer");
        lab = label_bufs(e->d_pcseed, e->nblocks);
        labels = &lab;
    }
    const PredState* pred = nullptr;
    if (e->params.policy == ASTCG_POLICY_LEARNED) {
        auto it = e->preds.find((int)(e->bx * 16 + e->by));
        if (it == e->preds.end()) return fail(e, ASTCG_ERR_NOT_CONFIGURED, "no predictor for this footprint");
        pred = &it->second;
    }
    if (pred) {
        e->fb_total = 0;
        err = launch_fb_reset(e->d_fb, s);
        if (err != hipSuccess) return hip_fail(e, err, "fallback counter");
    }
    return run_encode(e, *e->cur, A, (int)e->params.max_partitions, e->params.chunk_blocks, e->params.flags, s,
                      labels, pred, nullptr, e->params.policy == ASTCG_POLICY_EXHAUSTIVE ? kPartitionings : 0);
}

astcg_status astcg_encode_device_noalloc(astcg_encoder* e, const uint8_t* d_rgba, uint8_t* d_out, void* stream) {
    if (!e || !d_rgba || !d_out) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "null pointer");
    if (!e->configured) return fail(e, ASTCG_ERR_NOT_CONFIGURED, "encode before configure");
    if (!e->scratch) return fail(e, ASTCG_ERR_DEVICE, "no static scratch on this device");
    if (e->params.policy == ASTCG_POLICY_LEARNED)
        return fail(e, ASTCG_ERR_UNSUPPORTED, "the no-allocation encode runs the COMPAT and EXHAUSTIVE policies only");
    if (e->params.flags & (ASTCG_FLAG_PROFILE | ASTCG_FLAG_LABELS | ASTCG_FLAG_GRAPH))
        return fail(e, ASTCG_ERR_UNSUPPORTED,
                    "the no-allocation encode takes no PROFILE, LABELS or GRAPH flag (they need allocations)");
    const FpState& F = *e->cur;
    const Footprint& f = *F.fp;
    // chunk: as many blocks as fit in the scratch with their SymBest records
    size_t offs[13];
    const size_t fixed = chunk_bytes_per_block(f, offs, 0);
    const size_t per_block = ((chunk_bytes_per_block(f, offs, 1 << 16) - fixed) >> 16) + sizeof(SymBest);
    // exact fit: the sub-arrays' 256-byte alignment and the SymBest run
    // after them included
    auto need = [&](size_t n) {
        return ((chunk_bytes_per_block(f, offs, (int)n) + 255) & ~size_t(255)) + n * sizeof(SymBest);
    };
    size_t fit = kStaticScratchBytes > fixed ? (kStaticScratchBytes - fixed) / per_block + 1 : 0;
    while (fit > 0 && need(fit) > kStaticScratchBytes) --fit;
    if (e->params.chunk_blocks > fit)
        return fail(e, ASTCG_ERR_INVALID_ARGUMENT,
                    "chunk_blocks " + std::to_string(e->params.chunk_blocks) + " exceeds what the " +
                        std::to_string(kStaticScratchBytes >> 20) + " MiB static scratch holds at this footprint (" +
                        std::to_string(fit) + " blocks)");
    const int chunk = (int)std::min<size_t>(e->params.chunk_blocks ? e->params.chunk_blocks : fit,
                                            std::min<size_t>(e->nblocks, (size_t)chunk_size(f, e->nblocks, 0, nullptr)));
    if (chunk < 1) return fail(e, ASTCG_ERR_UNSUPPORTED, "footprint working set exceeds the static scratch");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    uint8_t* base = static_cast<uint8_t*>(e->scratch);
    if (need((size_t)chunk) > kStaticScratchBytes) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "scratch layout");
    const size_t used = chunk_bytes_per_block(f, offs, chunk);
    const ChunkBuffers B = chunk_buffers_at(base, F, offs);
    SymBest* best = reinterpret_cast<SymBest*>(base + ((used + 255) & ~size_t(255)));
    // the scratch keeps the previous call's contents: the per-mode counters
    // must start at zero (the search clears them again after each use)
    err = launch_zero_ints(B.bin_count, 2 * f.nbm * kRefineClasses + 2, s);
    if (err != hipSuccess) return hip_fail(e, err, "scratch clear");
    // each chunk is encoded as its own launch range: the texture's block
    // numbering starts at -blk0, so the chunk's blocks 0..n-1 are the image's
    // blocks blk0..blk0+n-1 (in-batch indices and output positions are the
    // texture's), and the chunk's SymBest records are packed before the next
    // chunk reuses them
    for (uint32_t blk0 = 0; blk0 < e->nblocks; blk0 += (uint32_t)chunk) {
        const int n = (int)std::min<uint32_t>((uint32_t)chunk, e->nblocks - blk0);
        ImageArgs A{};
        A.t0.rgba = d_rgba;
        A.t0.out = d_out;
        A.t0.width = (int)e->width;
        A.t0.height = (int)e->height;
        A.t0.blocks_x = (int)e->blocks_x;
        A.t0.first_block = -(int)blk0;
        A.t0.num_blocks = (int)e->nblocks;
        A.ntex = 1;
        A.num_blocks = n;
        A.blk0 = 0;
        A.nb = n;
        err = launch_chunk(F.T, A, B, best, (int)e->params.max_partitions, s, nullptr, nullptr, nullptr,
                           (e->params.flags & ASTCG_FLAG_DUAL_PLANE) != 0,
                           e->params.policy == ASTCG_POLICY_EXHAUSTIVE ? kPartitionings : 0);
        if (err != hipSuccess) return hip_fail(e, err, "search kernels");
        err = launch_pack(F.T, A, best, s, nullptr, (e->params.flags & ASTCG_FLAG_VOID_EXTENT) != 0);
        if (err != hipSuccess) return hip_fail(e, err, "pack kernel");
    }
    return ASTCG_OK;
}

astcg_status astcg_encode(astcg_encoder* e, const uint8_t* rgba, size_t rgba_len, uint8_t* out, size_t out_len) {
    if (!e || !rgba || !out) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "null pointer");
    if (!e->configured) return fail(e, ASTCG_ERR_NOT_CONFIGURED, "encode before configure");
    const size_t need_in = (size_t)e->width * e->height * 4, need_out = (size_t)e->nblocks * 16;
    if (rgba_len < need_in || out_len < need_out) return fail(e, ASTCG_ERR_BUFFER_TOO_SMALL, "buffer too small");
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    err = grow(reinterpret_cast<void**>(&e->d_rgba), &e->rgba_cap, need_in);
    if (err == hipSuccess) err = grow(reinterpret_cast<void**>(&e->d_out), &e->out_cap, need_out);
    if (err != hipSuccess) return hip_fail(e, err, "staging buffers");
    err = hipMemcpyAsync(e->d_rgba, rgba, need_in, hipMemcpyHostToDevice, e->stream);
    if (err != hipSuccess) return hip_fail(e, err, "upload");
    astcg_status st = astcg_encode_device(e, e->d_rgba, e->d_out, e->stream);
    if (st != ASTCG_OK) return st;
    err = hipMemcpyAsync(out, e->d_out, need_out, hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) return hip_fail(e, err, "download");
    return ASTCG_OK;
}

namespace {

// Checks one texture of a batch; returns its block count in *nb.
astcg_status check_texture(astcg_encoder* e, const astcg_texture& t, uint32_t i, uint64_t* nb) {
    char msg[96];
    if (!t.rgba || !t.out) {
        std::snprintf(msg, sizeof(msg), "texture %u: null pointer", i);
        return fail(e, ASTCG_ERR_INVALID_ARGUMENT, msg);
    }
    if (t.width == 0 || t.height == 0 || t.width >= (1u << 24) || t.height >= (1u << 24)) {
        std::snprintf(msg, sizeof(msg), "texture %u: dimensions must be in 1..2^24-1", i);
        return fail(e, ASTCG_ERR_INVALID_ARGUMENT, msg);
    }
    if (!valid_footprint((int)t.block_x, (int)t.block_y)) {
        std::snprintf(msg, sizeof(msg), "texture %u: not a 2D ASTC footprint", i);
        return fail(e, ASTCG_ERR_INVALID_FOOTPRINT, msg);
    }
    *nb = (uint64_t)((t.width + t.block_x - 1) / t.block_x) * ((t.height + t.block_y - 1) / t.block_y);
    if (t.out_len < *nb * 16) {
        std::snprintf(msg, sizeof(msg), "texture %u: out_len below %llu", i, (unsigned long long)(*nb * 16));
        return fail(e, ASTCG_ERR_BUFFER_TOO_SMALL, msg);
    }
    return ASTCG_OK;
}

// Encodes textures whose rgba/out are device pointers: one launch chain per
// footprint, over the textures of that footprint in their batch order.
astcg_status batch_device(astcg_encoder* e, const astcg_texture* tex, uint32_t count, const astcg_params& p,
                          hipStream_t s) {
    std::vector<int> keys;
    for (uint32_t i = 0; i < count; ++i) {
        const int key = (int)(tex[i].block_x * 16 + tex[i].block_y);
        if (std::find(keys.begin(), keys.end(), key) == keys.end()) keys.push_back(key);
    }
    // texture tables of all groups, uploaded once
    e->h_tex.clear();
    std::vector<size_t> group_start(keys.size() + 1, 0);
    std::vector<uint64_t> group_blocks(keys.size(), 0);
    for (size_t gi = 0; gi < keys.size(); ++gi) {
        const uint32_t bx = (uint32_t)keys[gi] / 16, by = (uint32_t)keys[gi] % 16;
        group_start[gi] = e->h_tex.size();
        uint64_t total = 0;
        for (uint32_t i = 0; i < count; ++i) {
            if (tex[i].block_x != bx || tex[i].block_y != by) continue;
            TexDesc d{};
            d.rgba = tex[i].rgba;
            d.out = tex[i].out;
            d.width = (int)tex[i].width;
            d.height = (int)tex[i].height;
            d.blocks_x = (int)((tex[i].width + bx - 1) / bx);
            d.num_blocks = d.blocks_x * (int)((tex[i].height + by - 1) / by);
            d.first_block = (int)total;
            total += (uint64_t)d.num_blocks;
            e->h_tex.push_back(d);
        }
        if (total > 0x7FFFFFFFull) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "too many blocks in one footprint group");
        group_blocks[gi] = total;
    }
    group_start[keys.size()] = e->h_tex.size();
    // fallback statistics over the whole batch
    e->fb_total = 0;
    if (p.policy == ASTCG_POLICY_LEARNED && e->d_fb) {
        hipError_t err = launch_fb_reset(e->d_fb, s);
        if (err != hipSuccess) return hip_fail(e, err, "fallback counter");
    }
    if (!e->h_tex.empty()) {
        const size_t bytes = sizeof(TexDesc) * e->h_tex.size();
        hipError_t err = grow(reinterpret_cast<void**>(&e->d_tex), &e->tex_cap, bytes);
        // the previous batch's table copy must have left the staging buffer
        if (err == hipSuccess && e->tex_copy_pending) err = hipEventSynchronize(e->tex_copied);
        if (err == hipSuccess) e->tex_copy_pending = false;
        if (err == hipSuccess && e->tex_pinned_cap < bytes) {
            if (e->h_tex_pinned) (void)hipHostFree(e->h_tex_pinned);
            e->h_tex_pinned = nullptr;
            e->tex_pinned_cap = 0;
            err = hipHostMalloc(reinterpret_cast<void**>(&e->h_tex_pinned), bytes, hipHostMallocDefault);
            if (err == hipSuccess) e->tex_pinned_cap = bytes;
        }
        if (err == hipSuccess && !e->tex_copied) err = hipEventCreateWithFlags(&e->tex_copied, hipEventDisableTiming);
        if (err == hipSuccess) {
            std::memcpy(e->h_tex_pinned, e->h_tex.data(), bytes);
            err = hipMemcpyAsync(e->d_tex, e->h_tex_pinned, bytes, hipMemcpyHostToDevice, s);
        }
        if (err == hipSuccess) err = hipEventRecord(e->tex_copied, s);
        if (err != hipSuccess) return hip_fail(e, err, "texture table");
        e->tex_copy_pending = true;
    }
    for (size_t gi = 0; gi < keys.size(); ++gi) {
        const FpState* F = nullptr;
        astcg_status st = get_footprint(e, (uint32_t)keys[gi] / 16, (uint32_t)keys[gi] % 16, &F);
        if (st != ASTCG_OK) return st;
        ImageArgs A{};
        A.ntex = (int)(group_start[gi + 1] - group_start[gi]);
        A.num_blocks = (int)group_blocks[gi];
        A.t0 = e->h_tex[group_start[gi]];
        A.tex = e->d_tex + group_start[gi];
        const PredState* pred = nullptr;
        std::vector<TexDesc> texs(e->h_tex.begin() + group_start[gi], e->h_tex.begin() + group_start[gi + 1]);
        if (p.policy == ASTCG_POLICY_LEARNED) {
            auto it = e->preds.find(keys[gi]);
            if (it == e->preds.end()) return fail(e, ASTCG_ERR_NOT_CONFIGURED, "no predictor for a batch footprint");
            pred = &it->second;
        }
        st = run_encode(e, *F, A, (int)p.max_partitions, p.chunk_blocks, p.flags, s, nullptr, pred, &texs,
                        p.policy == ASTCG_POLICY_EXHAUSTIVE ? kPartitionings : 0);
        if (st != ASTCG_OK) return st;
    }
    return ASTCG_OK;
}

}  // namespace

astcg_status astcg_encode_batch_device(astcg_encoder* e, const astcg_texture* textures, uint32_t count,
                                       const astcg_params* params, void* stream) {
    if (!e || (!textures && count)) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "null pointer");
    astcg_params p;
    astcg_status st = check_params(e, params, &p);
    if (st != ASTCG_OK) return st;
    for (uint32_t i = 0; i < count; ++i) {
        uint64_t nb = 0;
        st = check_texture(e, textures[i], i, &nb);
        if (st != ASTCG_OK) return st;
    }
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    // the search buffers may be laid out for another footprint afterwards;
    // a chunk cap of an earlier memory shortage is dropped
    e->layout_fp = nullptr;
    e->chunk_limit = 0;
    st = batch_device(e, textures, count, p, static_cast<hipStream_t>(stream));
    e->layout_fp = nullptr;
    return st;
}

astcg_status astcg_encode_batch(astcg_encoder* e, const astcg_texture* textures, uint32_t count,
                                const astcg_params* params) {
    if (!e || (!textures && count)) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "null pointer");
    astcg_params p;
    astcg_status st = check_params(e, params, &p);
    if (st != ASTCG_OK) return st;
    size_t in_bytes = 0, out_bytes = 0;
    std::vector<uint64_t> nbs(count);
    for (uint32_t i = 0; i < count; ++i) {
        st = check_texture(e, textures[i], i, &nbs[i]);
        if (st != ASTCG_OK) return st;
        in_bytes += ((size_t)textures[i].width * textures[i].height * 4 + 255) & ~size_t(255);
        out_bytes += ((size_t)nbs[i] * 16 + 255) & ~size_t(255);
    }
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    err = grow(reinterpret_cast<void**>(&e->d_rgba), &e->rgba_cap, in_bytes);
    if (err == hipSuccess) err = grow(reinterpret_cast<void**>(&e->d_out), &e->out_cap, out_bytes);
    if (err != hipSuccess) return hip_fail(e, err, "staging buffers");
    // stage every texture in device memory, then run the device path
    std::vector<astcg_texture> dev(textures, textures + count);
    size_t io = 0, oo = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const size_t n = (size_t)textures[i].width * textures[i].height * 4;
        err = hipMemcpyAsync(e->d_rgba + io, textures[i].rgba, n, hipMemcpyHostToDevice, e->stream);
        if (err != hipSuccess) return hip_fail(e, err, "upload");
        dev[i].rgba = e->d_rgba + io;
        dev[i].out = e->d_out + oo;
        dev[i].out_len = (size_t)nbs[i] * 16;
        io += (n + 255) & ~size_t(255);
        oo += ((size_t)nbs[i] * 16 + 255) & ~size_t(255);
    }
    e->layout_fp = nullptr;
    e->chunk_limit = 0;
    st = batch_device(e, dev.data(), count, p, e->stream);
    e->layout_fp = nullptr;
    if (st != ASTCG_OK) return st;
    for (uint32_t i = 0; i < count; ++i) {
        err = hipMemcpyAsync(textures[i].out, dev[i].out, (size_t)nbs[i] * 16, hipMemcpyDeviceToHost, e->stream);
        if (err != hipSuccess) return hip_fail(e, err, "download");
    }
    err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) return hip_fail(e, err, "download");
    return ASTCG_OK;
}

astcg_status astcg_get_symbolic(astcg_encoder* e, astcg_symbolic_block* out, size_t count) {
    if (!e || !out) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "null pointer");
    if (!e->configured) return fail(e, ASTCG_ERR_NOT_CONFIGURED, "not configured");
    if (count < e->nblocks) return fail(e, ASTCG_ERR_BUFFER_TOO_SMALL, "count below number of blocks");
    if (e->params.flags & ASTCG_FLAG_DUAL_PLANE)
        return fail(e, ASTCG_ERR_UNSUPPORTED, "SymbolicBlock has no plane-2 channel (ASTCG_FLAG_DUAL_PLANE)");
    std::vector<SymBest> h(e->nblocks);
    hipError_t err = hipSetDevice(e->device);
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err == hipSuccess) err = hipMemcpy(h.data(), e->d_best, sizeof(SymBest) * h.size(), hipMemcpyDeviceToHost);
    if (err != hipSuccess) return hip_fail(e, err, "symbolic download");
    for (size_t i = 0; i < h.size(); ++i) {
        astcg_symbolic_block& s = out[i];
        std::memset(&s, 0, sizeof(s));
        s.errorval = h[i].error;
        s.block_mode_index = h[i].raw_mode;
        s.partition_count = h[i].pc;
        s.partition_index = h[i].pc > 1 ? h[i].seed : 0;
        s.partition_formats_matched = h[i].matched;
        s.quant_mode = h[i].quant;
        for (int p = 0; p < 4; ++p) s.partition_formats[p] = (h[i].formats >> (4 * p)) & 0xF;
        for (int j = 0; j < 32; ++j) s.packed_color_values[j] = h[i].packed[j];
        for (int j = 0; j < 64; ++j) s.quantized_weights[j] = h[i].qw[j];
    }
    return ASTCG_OK;
}

astcg_status astcg_get_labels(astcg_encoder* e, int32_t* out, size_t count) {
    if (!e || !out) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "null pointer");
    if (!e->configured) return fail(e, ASTCG_ERR_NOT_CONFIGURED, "not configured");
    if (!(e->params.flags & ASTCG_FLAG_LABELS) || !e->d_pcseed)
        return fail(e, ASTCG_ERR_NOT_CONFIGURED, "encode with ASTCG_FLAG_LABELS first");
    if (count < e->nblocks) return fail(e, ASTCG_ERR_BUFFER_TOO_SMALL, "count below number of blocks");
    std::vector<SymBest> h(e->nblocks);
    std::vector<int16_t> ps(3 * (size_t)e->nblocks);
    const LabelBufs lab = label_bufs(e->d_pcseed, e->nblocks);
    hipError_t err = hipSetDevice(e->device);
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err == hipSuccess) err = hipMemcpy(h.data(), e->d_best, sizeof(SymBest) * h.size(), hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(ps.data(), lab.best_seed, sizeof(int16_t) * ps.size(), hipMemcpyDeviceToHost);
    if (err != hipSuccess) return hip_fail(e, err, "label download");
    const Footprint& f = *e->cur->fp;
    for (size_t i = 0; i < h.size(); ++i) {
        int32_t* o = out + 6 * i;
        o[0] = h[i].pc;
        o[1] = h[i].pc > 1 ? h[i].seed : 0;
        o[2] = h[i].raw_mode < 2048 ? f.raw_to_bm[h[i].raw_mode] : -1;
        for (int p = 0; p < 3; ++p) o[3 + p] = ps[(size_t)p * e->nblocks + i];
    }
    return ASTCG_OK;
}

astcg_status astcg_get_candidates(astcg_encoder* e, int16_t* seeds, float* errors, size_t count) {
    if (!e || !seeds || !errors) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "null pointer");
    if (!e->configured) return fail(e, ASTCG_ERR_NOT_CONFIGURED, "not configured");
    if (!(e->params.flags & ASTCG_FLAG_LABELS) || !e->d_pcseed)
        return fail(e, ASTCG_ERR_NOT_CONFIGURED, "encode with ASTCG_FLAG_LABELS first");
    if (count < e->nblocks) return fail(e, ASTCG_ERR_BUFFER_TOO_SMALL, "count below number of blocks");
    const size_t nb = e->nblocks;
    std::vector<int16_t> cs(12 * nb);
    std::vector<float> ce(12 * nb);
    const LabelBufs lab = label_bufs(e->d_pcseed, e->nblocks);
    hipError_t err = hipSetDevice(e->device);
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err == hipSuccess) err = hipMemcpy(cs.data(), lab.cand_seed, sizeof(int16_t) * cs.size(), hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(ce.data(), lab.cand_err, sizeof(float) * ce.size(), hipMemcpyDeviceToHost);
    if (err != hipSuccess) return hip_fail(e, err, "candidate download");
    for (size_t i = 0; i < nb; ++i)
        for (int j = 0; j < 12; ++j) {
            seeds[i * 12 + j] = cs[(size_t)j * nb + i];
            errors[i * 12 + j] = ce[(size_t)j * nb + i];
        }
    return ASTCG_OK;
}

void astcg_default_learned_params(astcg_learned_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->seeds = 0;
    p->modes = 16;
    p->slots = 5;
    p->fallback_below = 0.5f;
    p->count_below = 0.0f;
}

astcg_status astcg_set_predictor(astcg_encoder* e, uint32_t bx, uint32_t by, uint32_t hidden, const float* weights,
                                 size_t count, const astcg_learned_params* lp) {
    if (!e || !weights || hidden == 0 || hidden > 4096) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "bad predictor");
    if (!valid_footprint((int)bx, (int)by)) return fail(e, ASTCG_ERR_INVALID_FOOTPRINT, "not a 2D ASTC footprint");
    astcg_learned_params q;
    astcg_default_learned_params(&q);
    if (lp) q = *lp;
    if (q.seeds > 64 || q.modes < 1 || q.modes > 64 || q.slots < 1 || q.slots > 8 || q.partitionings > 4 ||
        q.refine_partitionings > 4 || !(q.early_out >= 0.0f) || !(q.full_search_above >= 0.0f) || !(q.count_above >= 0.0f) || q.full_search_modes > 1 || q.mismatch_candidates > (uint32_t)kPartitionings || !(q.fallback_above >= 0.0f) || q.rounds > 6 ||
        q.partitioned_slots > 8 || q.candidates > (uint32_t)kPartitionings ||
        ((q.candidates >= 1 || q.seeds != 0) && q.partitionings >= 1 && q.partitionings < 4))
        return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "learned params out of range");
    if (q.partitionings >= 1 && q.partitionings < 4 && hidden % 8 != 0)
        return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "ranking partitionings needs a hidden width divisible by 8");
    const Footprint* f = footprint((int)bx, (int)by);
    const int nin = 3 * (int)(bx * by), H = (int)hidden, nout = 4 + 3 * 1024 + f->nbm;
    const size_t sizes[6] = {(size_t)nin * H, (size_t)H, (size_t)H * H, (size_t)H, (size_t)H * nout, (size_t)nout};
    size_t need = 0;
    for (size_t z : sizes) need += z;
    if (count != need) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "predictor weight count mismatch");
    for (size_t i = 0; i < count; ++i)
        if (!std::isfinite(weights[i])) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "predictor weights must be finite");
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    const int key = (int)(bx * 16 + by);
    // captured launch chains hold the old predictor's buffers: drop them all
    if (!e->graphs.empty()) {
        (void)hipDeviceSynchronize();
        for (auto& g : e->graphs) (void)hipGraphExecDestroy(g.second);
        e->graphs.clear();
    }
    auto old = e->preds.find(key);
    if (old != e->preds.end()) {
        (void)hipDeviceSynchronize();
        free_pred(old->second);
        e->preds.erase(old);
    }
    if (!e->d_fb) {
        err = hipMalloc(reinterpret_cast<void**>(&e->d_fb), sizeof(unsigned long long));
        if (err == hipSuccess) err = hipMemset(e->d_fb, 0, sizeof(unsigned long long));
        if (err != hipSuccess) return hip_fail(e, err, "fallback counter");
    }
    PredState P;
    P.nin = nin;
    P.hidden = H;
    P.nout = nout;
    P.nmodes = f->nbm;
    P.lp = q;
    float* tmp = nullptr;
    err = hipMalloc(&tmp, sizes[4] * sizeof(float));
    size_t off = 0;
    for (int i = 0; i < 3 && err == hipSuccess; ++i) {
        const size_t wn = sizes[2 * i], bn = sizes[2 * i + 1];
        err = hipMalloc(&P.w16[i], wn * 2);
        if (err == hipSuccess) err = hipMalloc(&P.bias[i], bn * sizeof(float));
        if (err == hipSuccess) err = hipMemcpy(tmp, weights + off, wn * sizeof(float), hipMemcpyHostToDevice);
        if (err == hipSuccess && astcg_mlp_to_bf16(tmp, P.w16[i], (long long)wn, nullptr) != ASTCG_OK) err = hipErrorUnknown;
        if (err == hipSuccess) err = hipDeviceSynchronize();
        if (err == hipSuccess) err = hipMemcpy(P.bias[i], weights + off + wn, bn * sizeof(float), hipMemcpyHostToDevice);
        off += wn + bn;
    }
    const int sk = (int)q.seeds;
    const int head[15] = {0, 4, 4 + 1024, 4 + 2048, 4 + 3072, 4, 1024, 1024, 1024, f->nbm,
                          4, sk, sk, sk, (int)std::min<uint32_t>(q.modes, (uint32_t)f->nbm)};
    if (err == hipSuccess) err = hipMalloc(&P.d_head, sizeof(head));
    if (err == hipSuccess) err = hipMemcpy(P.d_head, head, sizeof(head), hipMemcpyHostToDevice);
    const bool fused_ok = H == 256 && nin <= 256 && 4 + f->nbm <= 256;
    if (err == hipSuccess && ((q.seeds == 0 && q.partitionings >= 1 && q.partitionings < 4) || (q.seeds > 0 && fused_ok))) {
        // seed-head columns of W3 transposed ([3 * 1024][H], one contiguous
        // row per seed) for ranking the compat positions, or for the seed
        // heads' top-k of the blocks that search a count from their seeds
        const float* w3 = weights + sizes[0] + sizes[1] + sizes[2] + sizes[3];
        std::vector<float> t((size_t)3 * 1024 * H);
        for (int c = 0; c < 3 * 1024; ++c)
            for (int j = 0; j < H; ++j) t[(size_t)c * H + j] = w3[(size_t)j * nout + 4 + c];
        err = hipMalloc(&P.w3_seed, t.size() * 2);
        if (err == hipSuccess) err = hipMemcpy(tmp, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice);
        if (err == hipSuccess && astcg_mlp_to_bf16(tmp, P.w3_seed, (long long)t.size(), nullptr) != ASTCG_OK)
            err = hipErrorUnknown;
        if (err == hipSuccess) err = hipDeviceSynchronize();
    }
    if (err == hipSuccess && fused_ok) {
        // fused-inference copies: W1^T [256][kin], W2^T, and the count and
        // mode columns of W3 as rows of [n3p][256] (zero padding)
        const int kin = (nin + 63) / 64 * 64, n3p = (4 + f->nbm + 15) / 16 * 16;
        const float* w1 = weights;
        const float* w2 = weights + sizes[0] + sizes[1];
        const float* w3 = w2 + sizes[2] + sizes[3];
        const float* b3 = w3 + sizes[4];
        std::vector<float> t1((size_t)256 * kin, 0.0f), t2((size_t)256 * 256), t3((size_t)n3p * 256, 0.0f), c3(n3p, 0.0f);
        for (int j = 0; j < 256; ++j)
            for (int i = 0; i < nin; ++i) t1[(size_t)j * kin + i] = w1[(size_t)i * 256 + j];
        for (int j = 0; j < 256; ++j)
            for (int i = 0; i < 256; ++i) t2[(size_t)j * 256 + i] = w2[(size_t)i * 256 + j];
        for (int c = 0; c < 4 + f->nbm; ++c) {
            const int src = c < 4 ? c : 4 + 3 * 1024 + (c - 4);
            for (int i = 0; i < 256; ++i) t3[(size_t)c * 256 + i] = w3[(size_t)i * nout + src];
            c3[c] = b3[src];
        }
        const std::vector<float>* ts[3] = {&t1, &t2, &t3};
        for (int i = 0; i < 3 && err == hipSuccess; ++i) {
            err = hipMalloc(&P.fw[i], ts[i]->size() * 2);
            if (err == hipSuccess) err = hipMemcpy(tmp, ts[i]->data(), ts[i]->size() * sizeof(float), hipMemcpyHostToDevice);
            if (err == hipSuccess && astcg_mlp_to_bf16(tmp, P.fw[i], (long long)ts[i]->size(), nullptr) != ASTCG_OK)
                err = hipErrorUnknown;
            if (err == hipSuccess) err = hipDeviceSynchronize();
        }
        if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&P.fb3), c3.size() * sizeof(float));
        if (err == hipSuccess) err = hipMemcpy(P.fb3, c3.data(), c3.size() * sizeof(float), hipMemcpyHostToDevice);
    }
    if (tmp) (void)hipFree(tmp);
    if (err != hipSuccess) {
        free_pred(P);
        return hip_fail(e, err, "predictor upload");
    }
    e->preds[key] = P;
    return ASTCG_OK;
}

astcg_status astcg_get_tuning(astcg_encoder* e, uint32_t bx, uint32_t by, astcg_tuning* out) {
    if (!e || !out) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "null pointer");
    if (!valid_footprint((int)bx, (int)by)) return fail(e, ASTCG_ERR_INVALID_FOOTPRINT, "not a 2D ASTC footprint");
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    FpState* F = nullptr;
    astcg_status st = get_footprint_mut(e, bx, by, &F);
    if (st != ASTCG_OK) return st;
    *out = tuning_of(*F);
    return ASTCG_OK;
}

astcg_status astcg_set_tuning(astcg_encoder* e, uint32_t bx, uint32_t by, const astcg_tuning* t) {
    if (!e || !t) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "null pointer");
    if (!valid_footprint((int)bx, (int)by)) return fail(e, ASTCG_ERR_INVALID_FOOTPRINT, "not a 2D ASTC footprint");
    if (t->ex_f4_max_tc > 64) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "ex_f4_max_tc above 64");
    if (t->refine_queue > 64) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "refine_queue above 64");
    if (t->refine_classes > 1) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "refine_classes must be 0 or 1");
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    FpState* F = nullptr;
    astcg_status st = get_footprint_mut(e, bx, by, &F);
    if (st != ASTCG_OK) return st;
    apply_tuning(*F, *t);
    return ASTCG_OK;
}

namespace {
// The tuning image: two colour gradients, block-scale noise and sharp-edged
// rectangles (xorshift64, fixed seed), so that every partition count and
// most trial modes are exercised.
void tune_image(int w, int h, std::vector<uint8_t>& px) {
    uint64_t s = 0x2545F4914F6CDD1Dull;
    auto next = [&s]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    };
    auto unit = [&next]() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); };
    double g[2][3], o[2][3], d[2][2];
    for (int k = 0; k < 2; ++k) {
        const double a = unit() * 6.283185307179586;
        d[k][0] = std::cos(a) / w;
        d[k][1] = std::sin(a) / h;
        for (int c = 0; c < 3; ++c) {
            g[k][c] = unit() * 160.0 - 80.0;
            o[k][c] = unit() * 100.0;
        }
    }
    px.assign((size_t)w * h * 4, 0);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            uint64_t hsh = (uint64_t)(x / 3) * 0x9E3779B97F4A7C15ull ^ (uint64_t)(y / 3) * 0xC2B2AE3D27D4EB4Full;
            hsh ^= hsh >> 29;
            for (int c = 0; c < 3; ++c) {
                double v = 40.0;
                for (int k = 0; k < 2; ++k) v += o[k][c] + g[k][c] * (x * d[k][0] + y * d[k][1]);
                v += (double)((hsh >> (8 * c)) & 15) - 7.5;
                px[((size_t)y * w + x) * 4 + c] = (uint8_t)std::min(255.0, std::max(0.0, v));
            }
            px[((size_t)y * w + x) * 4 + 3] = 255;
        }
    for (int r = 0; r < 40; ++r) {
        const int x0 = (int)(unit() * w), y0 = (int)(unit() * h);
        const int x1 = std::min(w, x0 + 4 + (int)(unit() * w / 6)), y1 = std::min(h, y0 + 4 + (int)(unit() * h / 6));
        uint8_t col[3];
        for (int c = 0; c < 3; ++c) col[c] = (uint8_t)(unit() * 255.0);
        for (int y = y0; y < y1; ++y)
            for (int x = x0; x < x1; ++x)
                for (int c = 0; c < 3; ++c) px[((size_t)y * w + x) * 4 + c] = col[c];
    }
}
}  // namespace

astcg_status astcg_autotune(astcg_encoder* e, uint32_t bx, uint32_t by, const char* cache_path, astcg_tuning* out) {
    if (!e) return ASTCG_ERR_INVALID_ARGUMENT;
    if (!valid_footprint((int)bx, (int)by)) return fail(e, ASTCG_ERR_INVALID_FOOTPRINT, "not a 2D ASTC footprint");
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    FpState* F = nullptr;
    astcg_status st = get_footprint_mut(e, bx, by, &F);
    if (st != ASTCG_OK) return st;
    const std::string dev = tune_device_id(e->device);
    astcg_tuning t;
    if (cache_path && tune_cache_lookup(cache_path, dev, bx, by, &t)) {
        apply_tuning(*F, t);
        if (out) *out = t;
        return ASTCG_OK;
    }
    // the caller's configuration, restored afterwards
    const bool was_configured = e->configured;
    const uint32_t sw = e->width, sh = e->height, sbx = e->bx, sby = e->by;
    const astcg_params sp = e->params;
    const int W = 1024, H = 1024;
    std::vector<uint8_t> img;
    tune_image(W, H, img);
    const size_t nb = (size_t)((W + bx - 1) / bx) * ((H + by - 1) / by);
    uint8_t *d_in = nullptr, *d_out = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    err = hipMalloc(reinterpret_cast<void**>(&d_in), img.size());
    if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&d_out), nb * 16);
    if (err == hipSuccess) err = hipMemcpy(d_in, img.data(), img.size(), hipMemcpyHostToDevice);
    for (int i = 0; i < 2 && err == hipSuccess; ++i) err = hipEventCreate(&ev[i]);
    auto time_encode = [&](uint32_t policy, float* ms) -> astcg_status {
        astcg_params p;
        astcg_default_params(&p);
        p.policy = policy;
        astcg_status s2 = astcg_configure(e, W, H, bx, by, &p);
        if (s2 == ASTCG_OK) s2 = astcg_encode_device(e, d_in, d_out, e->stream);
        hipError_t e2 = hipSuccess;
        if (s2 == ASTCG_OK) e2 = hipEventRecord(ev[0], e->stream);
        if (s2 == ASTCG_OK && e2 == hipSuccess) s2 = astcg_encode_device(e, d_in, d_out, e->stream);
        if (s2 == ASTCG_OK && e2 == hipSuccess) e2 = hipEventRecord(ev[1], e->stream);
        if (s2 == ASTCG_OK && e2 == hipSuccess) e2 = hipEventSynchronize(ev[1]);
        if (s2 == ASTCG_OK && e2 == hipSuccess) e2 = hipEventElapsedTime(ms, ev[0], ev[1]);
        if (s2 == ASTCG_OK && e2 != hipSuccess) s2 = hip_fail(e, e2, "autotune timing");
        return s2;
    };
    t = tuning_of(*F);
    st = err == hipSuccess ? ASTCG_OK : hip_fail(e, err, "autotune buffers");
    // each setting in turn, under the search whose dominant kernel it shapes
    struct Knob {
        uint32_t* field;
        int nchoices;
        uint32_t choices[4];
        uint32_t policy;
    };
    Knob knobs[4] = {{&t.refine_stage_max_tc, 2, {0u, 1u << 20, 0u, 0u}, ASTCG_POLICY_COMPAT},
                     {&t.refine_classes, 2, {0u, 1u, 0u, 0u}, ASTCG_POLICY_COMPAT},
                     {&t.refine_queue, 3, {0u, 16u, 64u, 0u}, ASTCG_POLICY_COMPAT},
                     {&t.ex_f4_max_tc, 2, {0u, 64u, 0u, 0u}, ASTCG_POLICY_EXHAUSTIVE}};
    for (int kn = 0; kn < 4 && st == ASTCG_OK; ++kn) {
        if (knobs[kn].field == &t.ex_f4_max_tc && F->fp->texel_count > 64) {
            t.ex_f4_max_tc = 0;  // float staging only up to 64 texels
            continue;
        }
        float best_ms = 0.0f;
        uint32_t best = *knobs[kn].field;
        for (int c = 0; c < knobs[kn].nchoices && st == ASTCG_OK; ++c) {
            *knobs[kn].field = knobs[kn].choices[c];
            apply_tuning(*F, t);
            float ms = 0.0f;
            st = time_encode(knobs[kn].policy, &ms);
            if (st == ASTCG_OK && (c == 0 || ms < best_ms)) {
                best_ms = ms;
                best = knobs[kn].choices[c];
            }
        }
        *knobs[kn].field = best;
        apply_tuning(*F, t);
    }
    for (int i = 0; i < 2; ++i)
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    if (d_in) (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    if (st != ASTCG_OK) return st;
    if (cache_path) {
        FILE* f = std::fopen(cache_path, "a");
        if (!f) return fail(e, ASTCG_ERR_IO, "cannot append to the tuning cache");
        std::fprintf(f, "astcg-tune 2 %s %ux%u %u %u %u %u %u\n", dev.c_str(), bx, by, t.refine_stage_max_tc,
                     t.ex_f4_max_tc, t.exw_max_list, t.refine_queue, t.refine_classes);
        const bool ok = std::fclose(f) == 0;
        if (!ok) return fail(e, ASTCG_ERR_IO, "cannot write the tuning cache");
    }
    if (out) *out = t;
    e->configured = false;
    if (was_configured) return astcg_configure(e, sw, sh, sbx, sby, &sp);
    return ASTCG_OK;
}

astcg_status astcg_fallback_rate(astcg_encoder* e, double* rate) {
    if (!e || !rate) return ASTCG_ERR_INVALID_ARGUMENT;
    *rate = 0.0;
    if (!e->fb_total || !e->d_fb) return ASTCG_OK;
    unsigned long long n = 0;
    hipError_t err = hipSetDevice(e->device);
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err == hipSuccess) err = hipMemcpy(&n, e->d_fb, sizeof(n), hipMemcpyDeviceToHost);
    if (err != hipSuccess) return hip_fail(e, err, "fallback counter");
    *rate = (double)n / (double)e->fb_total;
    return ASTCG_OK;
}

astcg_status astcg_trial_modes(uint32_t bx, uint32_t by, uint32_t* count, uint16_t* raw_modes, size_t cap) {
    if (!count) return ASTCG_ERR_INVALID_ARGUMENT;
    if (!valid_footprint((int)bx, (int)by)) return ASTCG_ERR_INVALID_FOOTPRINT;
    const Footprint* f = footprint((int)bx, (int)by);
    *count = (uint32_t)f->nbm;
    if (raw_modes)
        for (size_t i = 0; i < (size_t)f->nbm && i < cap; ++i) raw_modes[i] = f->bm[i].raw_mode;
    return ASTCG_OK;
}

astcg_status astcg_block_counts(const astcg_encoder* e, uint32_t* bxn, uint32_t* byn, uint32_t* nb) {
    if (!e) return ASTCG_ERR_INVALID_ARGUMENT;
    if (!e->configured) return ASTCG_ERR_NOT_CONFIGURED;
    if (bxn) *bxn = e->blocks_x;
    if (byn) *byn = e->blocks_y;
    if (nb) *nb = e->nblocks;
    return ASTCG_OK;
}

astcg_status astcg_symbolic_to_physical(uint32_t bx, uint32_t by, const astcg_symbolic_block* sb, uint8_t out[16]) {
    if (!sb || !out) return ASTCG_ERR_INVALID_ARGUMENT;
    return (astcg_status)symbolic_to_physical_host((int)bx, (int)by, *sb, out);
}

astcg_status astcg_encode_ise(uint32_t quant, uint32_t count, const uint8_t* in, uint8_t* out, uint32_t bit_offset) {
    if (!in || !out || quant > 20 || count > 64) return ASTCG_ERR_INVALID_ARGUMENT;
    static const int levels[21] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256};
    for (uint32_t i = 0; i < count; ++i)
        if (in[i] >= levels[quant]) return ASTCG_ERR_INVALID_ARGUMENT;
    encode_ise_host((int)quant, (int)count, in, out, (int)bit_offset);
    return ASTCG_OK;
}

uint32_t astcg_ise_bitcount(uint32_t count, uint32_t quant) { return (uint32_t)ise_bitcount((int)count, (int)quant); }

astcg_status astcg_astc_file_in_memory(uint32_t bx, uint32_t by, uint32_t w, uint32_t h, const uint8_t* data,
                                       size_t data_len, uint8_t* out, size_t out_cap, size_t* out_len) {
    if (!data || !out || !out_len) return ASTCG_ERR_INVALID_ARGUMENT;
    if (out_cap < data_len + 16) return ASTCG_ERR_BUFFER_TOO_SMALL;
    const uint32_t magic = 0x5CA1AB13u;
    uint8_t hdr[16] = {(uint8_t)(magic & 0xFF), (uint8_t)((magic >> 8) & 0xFF), (uint8_t)((magic >> 16) & 0xFF),
                       (uint8_t)(magic >> 24), (uint8_t)bx, (uint8_t)by, 1,
                       (uint8_t)(w & 0xFF), (uint8_t)((w >> 8) & 0xFF), (uint8_t)((w >> 16) & 0xFF),
                       (uint8_t)(h & 0xFF), (uint8_t)((h >> 8) & 0xFF), (uint8_t)((h >> 16) & 0xFF), 1, 0, 0};
    std::memcpy(out, hdr, 16);
    std::memcpy(out + 16, data, data_len);
    *out_len = data_len + 16;
    return ASTCG_OK;
}

astcg_status astcg_store_image(uint32_t bx, uint32_t by, uint32_t w, uint32_t h, const uint8_t* data, size_t data_len,
                               const char* path) {
    if (!data || !path) return ASTCG_ERR_INVALID_ARGUMENT;
    std::vector<uint8_t> buf(data_len + 16);
    size_t n = 0;
    astcg_status s = astcg_astc_file_in_memory(bx, by, w, h, data, data_len, buf.data(), buf.size(), &n);
    if (s != ASTCG_OK) return s;
    std::ofstream f(path, std::ios::out | std::ios::binary);
    if (!f) return ASTCG_ERR_IO;
    f.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)n);
    return f ? ASTCG_OK : ASTCG_ERR_IO;
}

astcg_status astcg_decode_image(uint32_t bx, uint32_t by, uint32_t w, uint32_t h, const uint8_t* blocks,
                                size_t blocks_len, uint8_t* rgba, size_t rgba_len) {
    if (!blocks || !rgba || w == 0 || h == 0) return ASTCG_ERR_INVALID_ARGUMENT;
    if (!valid_footprint((int)bx, (int)by)) return ASTCG_ERR_INVALID_FOOTPRINT;
    const uint32_t bxn = (w + bx - 1) / bx, byn = (h + by - 1) / by;
    if (blocks_len < (size_t)bxn * byn * 16 || rgba_len < (size_t)w * h * 4) return ASTCG_ERR_BUFFER_TOO_SMALL;
    std::vector<uint8_t> tex(bx * by * 4);
    astcg_status worst = ASTCG_OK;
    for (uint32_t j = 0; j < byn; ++j)
        for (uint32_t i = 0; i < bxn; ++i) {
            const int r = decode_block_rgba8((int)bx, (int)by, blocks + ((size_t)j * bxn + i) * 16, tex.data());
            if (r != ASTCG_OK) worst = ASTCG_ERR_UNSUPPORTED;
            for (uint32_t y = 0; y < by; ++y)
                for (uint32_t x = 0; x < bx; ++x) {
                    const uint32_t px = i * bx + x, py = j * by + y;
                    if (px >= w || py >= h) continue;
                    std::memcpy(rgba + ((size_t)py * w + px) * 4, &tex[(y * bx + x) * 4], 4);
                }
        }
    return worst;
}

astcg_status astcg_decode_device(astcg_encoder* e, uint32_t bx, uint32_t by, uint32_t w, uint32_t h,
                                 const uint8_t* d_blocks, uint8_t* d_rgba, void* stream) {
    if (!e || !d_blocks || !d_rgba || w == 0 || h == 0 || w >= (1u << 24) || h >= (1u << 24))
        return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "bad decode arguments");
    if (!valid_footprint((int)bx, (int)by)) return fail(e, ASTCG_ERR_INVALID_FOOTPRINT, "not a 2D ASTC footprint");
    const uint64_t nb = (uint64_t)((w + bx - 1) / bx) * ((h + by - 1) / by);
    if (nb > 0x7FFFFFFFull) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "too many blocks");
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    const FpState* F = nullptr;
    astcg_status st = get_footprint(e, bx, by, &F);
    if (st != ASTCG_OK) return st;
    err = launch_decode(F->T, d_blocks, (int)w, (int)h, d_rgba, static_cast<hipStream_t>(stream));
    return err == hipSuccess ? ASTCG_OK : hip_fail(e, err, "decode kernel");
}

astcg_status astcg_decode(astcg_encoder* e, uint32_t bx, uint32_t by, uint32_t w, uint32_t h, const uint8_t* blocks,
                          size_t blocks_len, uint8_t* rgba, size_t rgba_len) {
    if (!e || !blocks || !rgba || w == 0 || h == 0) return fail(e, ASTCG_ERR_INVALID_ARGUMENT, "bad decode arguments");
    if (!valid_footprint((int)bx, (int)by)) return fail(e, ASTCG_ERR_INVALID_FOOTPRINT, "not a 2D ASTC footprint");
    const size_t nb = (size_t)((w + bx - 1) / bx) * ((h + by - 1) / by);
    if (blocks_len < nb * 16 || rgba_len < (size_t)w * h * 4) return fail(e, ASTCG_ERR_BUFFER_TOO_SMALL, "buffer too small");
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return hip_fail(e, err, "hipSetDevice");
    // the staging buffers double as decode input (blocks) and output (texels)
    err = grow(reinterpret_cast<void**>(&e->d_out), &e->out_cap, nb * 16);
    if (err == hipSuccess) err = grow(reinterpret_cast<void**>(&e->d_rgba), &e->rgba_cap, (size_t)w * h * 4);
    if (err == hipSuccess) err = hipMemcpyAsync(e->d_out, blocks, nb * 16, hipMemcpyHostToDevice, e->stream);
    if (err != hipSuccess) return hip_fail(e, err, "decode staging");
    astcg_status st = astcg_decode_device(e, bx, by, w, h, e->d_out, e->d_rgba, e->stream);
    if (st != ASTCG_OK) return st;
    err = hipMemcpyAsync(rgba, e->d_rgba, (size_t)w * h * 4, hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    return err == hipSuccess ? ASTCG_OK : hip_fail(e, err, "decode download");
}

astcg_status astcg_physical_to_symbolic(uint32_t bx, uint32_t by, const uint8_t in[16], astcg_symbolic_block* out) {
    if (!in || !out) return ASTCG_ERR_INVALID_ARGUMENT;
    return (astcg_status)physical_to_symbolic_host((int)bx, (int)by, in, out);
}

astcg_status astcg_kernel_times(astcg_encoder* e, astcg_kernel_time* out, size_t cap, size_t* count) {
    if (!e || !count) return ASTCG_ERR_INVALID_ARGUMENT;
    if (!e->prof) return fail(e, ASTCG_ERR_NOT_CONFIGURED, "profiling not enabled");
    (void)hipSetDevice(e->device);
    e->prof->resolve();
    size_t n = 0;
    for (int k = 0; k < K_COUNT; ++k) {
        if (!e->prof->launches[k]) continue;
        if (out && n < cap) {
            std::snprintf(out[n].name, sizeof(out[n].name), "%s", kernel_name(k));
            out[n].total_ms = e->prof->total_ms[k];
            out[n].launches = (uint64_t)e->prof->launches[k];
        }
        ++n;
    }
    *count = n;
    return ASTCG_OK;
}

astcg_status astcg_reset_kernel_times(astcg_encoder* e) {
    if (!e) return ASTCG_ERR_INVALID_ARGUMENT;
    if (!e->prof) return fail(e, ASTCG_ERR_NOT_CONFIGURED, "profiling not enabled");
    (void)hipSetDevice(e->device);
    e->prof->resolve();
    for (int k = 0; k < K_COUNT; ++k) { e->prof->total_ms[k] = 0.0; e->prof->launches[k] = 0; }
    return ASTCG_OK;
}

const char* astcg_last_error(const astcg_encoder* e) { return e ? e->last_error.c_str() : "null encoder"; }

const char* astcg_status_string(astcg_status s) {
    switch (s) {
        case ASTCG_OK: return "ok";
        case ASTCG_ERR_INVALID_ARGUMENT: return "invalid argument";
        case ASTCG_ERR_INVALID_FOOTPRINT: return "invalid footprint";
        case ASTCG_ERR_BUFFER_TOO_SMALL: return "buffer too small";
        case ASTCG_ERR_NOT_CONFIGURED: return "not configured";
        case ASTCG_ERR_DEVICE: return "device error";
        case ASTCG_ERR_OUT_OF_MEMORY: return "out of memory";
        case ASTCG_ERR_IO: return "i/o error";
        case ASTCG_ERR_UNSUPPORTED: return "unsupported";
    }
    return "unknown";
}

}  // extern "C"
