// rowgroup.hip -- row-group decode: many columns, one persistent launch over a device-side work list.
//
// Replaces the per-column decode loop that sits behind Reader::get_chunk / Interpreter::Decoding in the reference
// (src/reader/reader.cpp:19-61, src/expression/interpreter.cpp:105-137) with a work list of (page, vector range)
// tiles decoded by one persistent kernel (k_rowgroup.hip); mode 1 keeps the one-launch-per-page form for
// comparison. The persistent launch can be captured in a HIP graph and replayed.
#include <algorithm>
#include <queue>
#include <cstring>
#include <string>
#include <memory>
#include <vector>

#include "../../include/fls_gpu.h"
#include "format.hpp"
#include "gpu_internal.hpp"
#include "launch.hpp"
#include "page_kernels.hpp"
#include "rowgroup.hpp"

using namespace fls;

// fls_gpu_plan internals used here (defined in gpu_api.hip)
namespace fls {
const Column& plan_column(const fls_gpu_plan* p);
int           plan_decode_dictionaries(fls_gpu_plan* p, const uint8_t* dev_encoded, hipStream_t s);
const void*   plan_dictionary(const fls_gpu_plan* p, size_t page);
}  // namespace fls

struct fls_gpu_rowgroup {
    std::vector<fls_gpu_plan*> plans;
    std::vector<uint64_t>      dev_off, out_off;
    std::vector<uint8_t>       types;
    std::vector<bool>          fused;  // column decoded by the persistent kernel
    uint64_t                   out_total = 0;
    std::vector<RgTile>        tiles;       // in launch order: one contiguous run per wave
    std::vector<uint32_t>      wave_begin;  // runs' bounds, waves + 1 entries
    RgTile*                    d_tiles      = nullptr;
    uint32_t*                  d_wave_begin = nullptr;
    uint32_t*                  d_err        = nullptr;
    unsigned                   wave_lds = 0, out_lds = 0, blocks = 0, waves = kRgWaves;  // waves per block
    // captured graph
    hipGraph_t     graph = nullptr;
    hipGraphExec_t exec  = nullptr;
    const uint8_t* g_enc = nullptr;
    uint8_t*       g_out = nullptr;
    // parallel branches of the per-page form
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t>  join;
    hipEvent_t               fork = nullptr;
    std::vector<size_t>      order;  // columns, largest first
};

namespace fls {

std::string rg_validate_tiles(const std::vector<RgTile>& tiles, const std::vector<RgPageExtent>& pages,
                              uint64_t enc_size, uint64_t out_size, unsigned wave_lds) {
    // buffer count and staged range count of each kind family: int UNC / FFOR / DELTA, RLE, dictionary codes
    // UNCOMPRESSED / FFOR / DELTA
    static const unsigned kBuffers[7] = {1, 3, 4, 5, 1, 3, 4};
    static const unsigned kRanges[7]  = {1, 1, 2, 3, 1, 1, 2};
    auto r16 = [](uint64_t n) { return (n + 15) & ~15ull; };
    for (size_t n = 0; n < tiles.size(); ++n) {
        const RgTile& t   = tiles[n];
        const auto    why = [&](const char* what) { return "tile " + std::to_string(n) + ": " + what; };
        if (t.kind >= RG_KIND_COUNT) return why("unknown kind");
        if (t.nv == 0 || t.nv > kRgMaxTileVec) return why("vector count outside 1..64");
        if ((uint64_t)t.v0 + t.nv > t.vec_c) return why("vector range outside its page");
        const RgPageExtent* pe = nullptr;
        for (const auto& e : pages)
            if (e.page == t.page) pe = &e;
        if (!pe) return why("page not in the row-group");
        if (pe->page + pe->page_size > enc_size) return why("page outside the encoded buffer");
        if (pe->page_size < 16) return why("page shorter than 16 bytes");  // the kernel's stand-in load address
        if (t.arr_c != kBuffers[t.kind / 4]) return why("buffer count does not match the kind");
        for (unsigned b = 0; b < t.arr_c; ++b)
            if (t.off[b] >= pe->page_size) return why("buffer offset outside its page");
        if (t.n_rng != kRanges[t.kind / 4]) return why("staged range count does not match the kind");
        uint64_t chunks = 0, lds_end = 0;
        for (unsigned r = 0; r < 3; ++r) {
            if (r >= t.n_rng && (t.rlen[r] || t.rsrc[r])) return why("unused staged range is not empty");
            if ((uint64_t)t.rsrc[r] + r16(t.rlen[r]) > pe->page_size) return why("staged range outside its page");
            if (t.rlds[r] % 16) return why("staged range not 16-byte aligned in LDS");
            if (r && r < t.n_rng && t.rlds[r] < lds_end) return why("staged ranges overlap in LDS");
            if (r < t.n_rng) lds_end = t.rlds[r] + r16(t.rlen[r]) + (r == 0 ? 128 : 0);  // range 0: + pad row
            chunks += (t.rlen[r] + 15) / 16;
        }
        if (chunks != t.chunks) return why("chunk count differs from the ranges");
        for (unsigned r = 0, first = 0; r < t.n_rng; first += (t.rlen[r] + 15) / 16, ++r)
            if ((int64_t)t.sdelta[r] != (int64_t)t.rsrc[r] - 16 * (int64_t)first ||
                (int64_t)t.ldelta[r] != (int64_t)t.rlds[r] - 16 * (int64_t)first)
                return why("chunk deltas differ from the ranges");
        if (lds_end + kRgDump > wave_lds) return why("staged ranges exceed the wave's LDS slot");
        const uint64_t vb = 128ull * rg_kind_bits(t.kind);  // bytes of one decoded vector
        if (t.out != pe->out || (uint64_t)t.vec_c * vb != pe->out_bytes) return why("output range differs from its page");
        if (t.out + (uint64_t)(t.v0 + t.nv) * vb > out_size) return why("output past the row-group's output");
        if (t.kind >= RG_DICT1_8 && (!t.dict || !t.dict_n)) return why("dictionary kind without a dictionary");
    }
    return "";
}

std::vector<uint32_t> rg_assign_waves(std::vector<RgTile>& tiles, uint32_t waves) {
    waves = std::max(1u, waves);
    // estimated work of a tile: its decoded and staged bytes plus a fixed cost per vector and per tile
    auto cost = [](const RgTile& t) {
        return (uint64_t)t.nv * (128ull * rg_kind_bits(t.kind) + 2048) + 16ull * t.chunks + 4096;
    };
    std::vector<uint32_t> by_cost(tiles.size());
    for (uint32_t i = 0; i < by_cost.size(); ++i) by_cost[i] = i;
    std::stable_sort(by_cost.begin(), by_cost.end(), [&](uint32_t a, uint32_t b) { return cost(tiles[a]) > cost(tiles[b]); });
    using Load = std::pair<uint64_t, uint32_t>;  // (work so far, wave)
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (uint32_t w = 0; w < waves; ++w) heap.push({0, w});
    std::vector<uint32_t> owner(tiles.size());
    for (uint32_t i : by_cost) {
        Load l = heap.top();
        heap.pop();
        owner[i] = l.second;
        l.first += cost(tiles[i]);
        heap.push(l);
    }
    std::vector<uint32_t> begin(waves + 1, 0);
    for (uint32_t i = 0; i < tiles.size(); ++i) ++begin[owner[i] + 1];
    for (uint32_t w = 0; w < waves; ++w) begin[w + 1] += begin[w];
    std::vector<RgTile>   ordered(tiles.size());
    std::vector<uint32_t> at(begin.begin(), begin.end() - 1);
    for (uint32_t i = 0; i < tiles.size(); ++i) ordered[at[owner[i]]++] = tiles[i];
    tiles.swap(ordered);
    return begin;
}

}  // namespace fls

namespace {

// Persistent-kernel kind of a page (-1: the page decodes through its per-page kernel). The kernel reads per-vector
// metadata at fixed strides, which is how the encoders lay pages out (PageEntry::fixed_meta).
int kind_of(const PageEntry& pe, unsigned W) {
    if (!pe.fixed_meta || !pe.vec_c) return -1;
    const int wc = W == 8 ? 0 : W == 16 ? 1 : W == 32 ? 2 : 3;
    switch (pe.exp_t) {
    case EXP_UNCOMPRESSED: return RG_UNC8 + wc;
    case EXP_FFOR: return pe.var == 1 ? RG_FFOR8 + wc : -1;
    case EXP_DELTA: return pe.var == 1 ? RG_DELTA8 + wc : -1;
    case EXP_RLE: return pe.var == 2 ? RG_RLE8 + wc : -1;
    case EXP_DICT:
    case EXP_DICT_FRQ:
    case EXP_DICT_VAL:
        if (!pe.dict_n) return -1;
        return pe.var == 1 ? RG_DICT1_8 + wc : pe.var == 2 ? RG_DICT2_8 + wc : pe.var == 7 ? RG_DICT7_8 + wc : -1;
    default: return -1;
    }
}

void free_rg(fls_gpu_rowgroup* rg) {
    for (auto* p : rg->plans) fls_gpu_plan_destroy(p);
    if (rg->d_tiles) (void)hipFree(rg->d_tiles);
    if (rg->d_wave_begin) (void)hipFree(rg->d_wave_begin);
    if (rg->d_err) (void)hipFree(rg->d_err);
    if (rg->exec) (void)hipGraphExecDestroy(rg->exec);
    if (rg->graph) (void)hipGraphDestroy(rg->graph);
    for (auto s : rg->branch) (void)hipStreamDestroy(s);
    for (auto ev : rg->join) (void)hipEventDestroy(ev);
    if (rg->fork) (void)hipEventDestroy(rg->fork);
    delete rg;
}

int decode_pages(fls_gpu_rowgroup* rg, const uint8_t* enc, uint8_t* out, hipStream_t s) {
    for (size_t c = 0; c < rg->plans.size(); ++c) {
        const uint64_t bytes = fls_gpu_plan_n_values(rg->plans[c]) * (dtype_bits(rg->types[c]) / 8);
        int rc = fls_gpu_plan_decode(rg->plans[c], enc + rg->dev_off[c], out + rg->out_off[c], bytes, s);
        if (rc) return rc;
    }
    return FLS_OK;
}

// Per-page launches spread over parallel branches: `cs` forks into rg->branches streams (columns assigned longest
// first to the least-loaded branch), which join back into `cs`. Inside a graph capture the branches become parallel
// graph paths; outside, they overlap the ramp-up and drain of small kernels.
int decode_pages_branched(fls_gpu_rowgroup* rg, const uint8_t* enc, uint8_t* out, hipStream_t cs) {
    const unsigned nb = (unsigned)rg->branch.size();
    hipError_t     e  = hipEventRecord(rg->fork, cs);
    for (unsigned b = 0; b < nb && e == hipSuccess; ++b) e = hipStreamWaitEvent(rg->branch[b], rg->fork, 0);
    if (e != hipSuccess) return hip_fail(e, "fork");
    std::vector<uint64_t> load(nb, 0);
    for (size_t c : rg->order) {
        const unsigned b = (unsigned)(std::min_element(load.begin(), load.end()) - load.begin());
        const uint64_t bytes = fls_gpu_plan_n_values(rg->plans[c]) * (dtype_bits(rg->types[c]) / 8);
        int rc = fls_gpu_plan_decode(rg->plans[c], enc + rg->dev_off[c], out + rg->out_off[c], bytes, rg->branch[b]);
        if (rc) return rc;
        load[b] += bytes + fls_gpu_plan_encoded_size(rg->plans[c]);
    }
    for (unsigned b = 0; b < nb && e == hipSuccess; ++b) {
        e = hipEventRecord(rg->join[b], rg->branch[b]);
        if (e == hipSuccess) e = hipStreamWaitEvent(cs, rg->join[b], 0);
    }
    return e == hipSuccess ? FLS_OK : hip_fail(e, "join");
}

// The persistent launch's work list for parsed columns (host copies `host[c]`): every page of a fused column (a
// column is fused when all its pages have a kind) is cut into tiles of consecutive vectors, each at most
// `tile_kib` KiB of decoded values, 64 vectors and kRgStage staged bytes (a single vector may exceed the last), with
// the byte ranges the kernel stages for it. Also the LDS bytes per wave the launch needs and the columns' output
// offsets. dict_of(c, i): device address of column c's page-i dictionary. The list is checked (rg_validate_tiles)
// before it is returned.
struct WorkList {
    std::vector<RgTile>   tiles;
    std::vector<bool>     fused;
    std::vector<uint64_t> out_off;
    uint64_t              out_total = 0;
    unsigned              wave_lds  = 0;
};

// Per-vector staged bytes of range 0..2 for a page's kind; packed rows are contiguous across vectors when every
// vector's entry point is the previous one's plus its rows (how the encoders write pages).
struct VecStage {
    uint32_t len[3];
};
bool vec_stage(const uint8_t* pg, const PageEntry& pe, int kind, uint32_t v, VecStage& vs) {
    const unsigned fam = kind / 4, W = 8u << (kind % 4);
    auto           ep  = [&](uint32_t vv, unsigned b) { return (uint32_t)ep_of(pg, pe.arr_c, vv, b); };
    auto           bwc = [&](unsigned b) { return (uint32_t)pg[pe.boff[b] + v]; };  // fixed-stride bit width
    vs = VecStage {};
    switch (fam) {
    case 0: vs.len[0] = 128 * W; return true;
    case 1: vs.len[0] = 128 * bwc(0); return v + 1 == pe.vec_c || ep(v + 1, 2) == ep(v, 2) + vs.len[0];
    case 2:
        vs.len[0] = 128 * bwc(1), vs.len[1] = 128;
        return v + 1 == pe.vec_c || ep(v + 1, 3) == ep(v, 3) + vs.len[0];
    case 3: {
        const uint32_t r0 = ep(v, 0), r1 = v + 1 < pe.vec_c ? ep(v + 1, 0) : pe.boff[1] - pe.boff[0];
        vs.len[0] = 128 * bwc(2), vs.len[1] = r1 - r0, vs.len[2] = 128;
        return r1 >= r0 && (v + 1 == pe.vec_c || ep(v + 1, 4) == ep(v, 4) + vs.len[0]);
    }
    case 4: vs.len[0] = 4096; return true;
    case 5: vs.len[0] = 128 * bwc(0); return v + 1 == pe.vec_c || ep(v + 1, 2) == ep(v, 2) + vs.len[0];
    default:
        vs.len[0] = 128 * bwc(1), vs.len[1] = 128;
        return v + 1 == pe.vec_c || ep(v + 1, 3) == ep(v, 3) + vs.len[0];
    }
}
// Start of staged range r of vector v (byte offset from the page start).
uint32_t range_src(const uint8_t* pg, const PageEntry& pe, int kind, uint32_t v, unsigned r) {
    const unsigned fam = kind / 4, W = 8u << (kind % 4);
    auto           ep  = [&](unsigned b) { return (uint32_t)ep_of(pg, pe.arr_c, v, b); };
    switch (fam) {
    case 0: return pe.boff[0] + v * 128 * W;
    case 1: return pe.boff[2] + ep(2);
    case 2: return r == 0 ? pe.boff[3] + ep(3) : pe.boff[0] + 128 * v;
    case 3: return r == 0 ? pe.boff[4] + ep(4) : r == 1 ? pe.boff[0] + ep(0) : pe.boff[1] + 128 * v;
    case 4: return pe.boff[0] + 4096 * v;
    case 5: return pe.boff[2] + ep(2);
    default: return r == 0 ? pe.boff[3] + ep(3) : pe.boff[0] + 128 * v;
    }
}

template <typename DictOf>
std::string plan_work_list(const std::vector<const Column*>& cols, const uint8_t* const* host, const uint8_t* types,
                           const uint64_t* device_offsets, const uint64_t* column_sizes, unsigned tile_kib,
                           DictOf&& dict_of, WorkList& wl) {
    static const unsigned     kRanges[7] = {1, 1, 2, 3, 1, 1, 2};
    auto                      r16        = [](uint64_t n) { return (uint32_t)((n + 15) & ~15ull); };
    std::vector<RgPageExtent> extents;
    uint64_t                  enc_size = 0;
    unsigned                  lds_need = 0;
    for (size_t c = 0; c < cols.size(); ++c) {
        const Column&  col   = *cols[c];
        const unsigned W     = dtype_bits(types[c]);
        const uint64_t bytes = col.n_values * (W / 8);
        enc_size             = std::max(enc_size, device_offsets[c] + column_sizes[c]);
        wl.out_off.push_back(wl.out_total);
        bool fused = true;
        for (const auto& pe : col.pages) fused &= kind_of(pe, W) >= 0;
        // every vector's staged ranges must be contiguous within its page
        std::vector<std::vector<VecStage>> stages(col.pages.size());
        for (size_t i = 0; fused && i < col.pages.size(); ++i) {
            const PageEntry& pe = col.pages[i];
            const int        kind = kind_of(pe, W);
            stages[i].resize(pe.vec_c);
            for (uint32_t v = 0; fused && v < pe.vec_c; ++v)
                fused = vec_stage(host[c] + pe.page_off, pe, kind, v, stages[i][v]);
        }
        wl.fused.push_back(fused);
        if (fused) {
            const unsigned cap_vec =
                std::max(1u, std::min(kRgMaxTileVec, (unsigned)((tile_kib * 1024ull) / (1024ull * (W / 8)))));
            for (size_t i = 0; i < col.pages.size(); ++i) {
                const PageEntry& pe   = col.pages[i];
                const uint8_t*   pg   = host[c] + pe.page_off;
                const int        kind = kind_of(pe, W);
                const unsigned   nr   = kRanges[kind / 4];
                extents.push_back({device_offsets[c] + pe.page_off, pe.page_size, wl.out_total + pe.out_index * (W / 8),
                                   (uint64_t)pe.vec_c * 1024 * (W / 8)});
                auto need = [&](const uint32_t (&l)[3]) { return r16(l[0]) + 128 + r16(l[1]) + r16(l[2]); };
                for (uint32_t v0 = 0; v0 < pe.vec_c;) {
                    uint32_t len[3] = {0, 0, 0}, nv = 0;
                    while (v0 + nv < pe.vec_c && nv < cap_vec) {
                        const VecStage& vs   = stages[i][v0 + nv];
                        const uint32_t  l[3] = {len[0] + vs.len[0], len[1] + vs.len[1], len[2] + vs.len[2]};
                        if (nv && need(l) > kRgStage) break;
                        len[0] = l[0], len[1] = l[1], len[2] = l[2], ++nv;
                    }
                    RgTile t {};
                    t.page   = device_offsets[c] + pe.page_off;
                    t.out    = wl.out_total + pe.out_index * (W / 8);
                    t.dict   = dict_of(c, i);
                    t.dict_n = pe.dict_n;
                    t.v0     = v0;
                    t.nv     = (uint16_t)nv;
                    t.vec_c  = pe.vec_c;
                    t.kind   = (uint16_t)kind;
                    t.arr_c  = (uint16_t)pe.arr_c;
                    t.n_rng  = (uint16_t)nr;
                    for (unsigned k = 0; k < 5 && k < pe.arr_c; ++k) t.off[k] = pe.boff[k];
                    uint32_t at = 0;
                    for (unsigned r = 0; r < nr; ++r) {
                        t.rsrc[r]   = range_src(pg, pe, kind, v0, r);
                        t.rlen[r]   = len[r];
                        t.rlds[r]   = at;
                        t.sdelta[r] = (int32_t)t.rsrc[r] - 16 * (int32_t)t.chunks;
                        t.ldelta[r] = (int32_t)t.rlds[r] - 16 * (int32_t)t.chunks;
                        at += r16(len[r]) + (r == 0 ? 128 : 0);  // range 0 is followed by its pad row
                        t.chunks += (len[r] + 15) / 16;
                    }
                    lds_need = std::max(lds_need, at);
                    wl.tiles.push_back(t);
                    v0 += nv;
                }
            }
        }
        wl.out_total += (bytes + 255) & ~255ull;
    }
    wl.wave_lds = wl.tiles.empty() ? 0 : lds_need + kRgDump;
    // widest values first, each kind's tiles together: the long tiles start early and the short 8-bit ones fill
    // the tail, and a compute unit's consecutive tiles mostly share one code path
    std::stable_sort(wl.tiles.begin(), wl.tiles.end(), [](const RgTile& a, const RgTile& b) {
        const unsigned wa = rg_kind_bits(a.kind), wb = rg_kind_bits(b.kind);
        return wa != wb ? wa > wb : a.kind < b.kind;
    });
    return rg_validate_tiles(wl.tiles, extents, enc_size, wl.out_total, wl.wave_lds);
}

unsigned tile_kib_tuned() {
    const Tune tn = tune_for(FAM_ROWGROUP, 0, 255);  // block = KiB of decoded values per tile
    return tn.block ? tn.block : 64;
}

// Plans with the tuned tile size, then with smaller tiles while the list has fewer than `min_tiles` tiles (a small
// row-group would otherwise leave most waves without work), down to 16 KiB per tile.
template <typename DictOf>
std::string plan_work_list_sized(const std::vector<const Column*>& cols, const uint8_t* const* host,
                                 const uint8_t* types, const uint64_t* device_offsets, const uint64_t* column_sizes,
                                 uint64_t min_tiles, DictOf&& dict_of, WorkList& wl) {
    unsigned tile_kib = tile_kib_tuned();
    for (;;) {
        wl = WorkList {};
        std::string bad = plan_work_list(cols, host, types, device_offsets, column_sizes, tile_kib, dict_of, wl);
        if (!bad.empty() || wl.tiles.size() >= min_tiles || tile_kib <= 16) return bad;
        tile_kib = std::max(16u, std::min(tile_kib / 2, (unsigned)(tile_kib * wl.tiles.size() / min_tiles)));
    }
}

int decode_fused(fls_gpu_rowgroup* rg, const uint8_t* enc, uint8_t* out, hipStream_t s) {
    for (size_t c = 0; c < rg->plans.size(); ++c) {
        if (!rg->fused[c]) {
            const uint64_t bytes = fls_gpu_plan_n_values(rg->plans[c]) * (dtype_bits(rg->types[c]) / 8);
            int rc = fls_gpu_plan_decode(rg->plans[c], enc + rg->dev_off[c], out + rg->out_off[c], bytes, s);
            if (rc) return rc;
        } else {
            int rc = plan_decode_dictionaries(rg->plans[c], enc + rg->dev_off[c], s);
            if (rc) return rc;
        }
    }
    if (rg->tiles.empty()) return FLS_OK;
    hipError_t e = launch_rowgroup(enc, out, rg->d_tiles, rg->d_wave_begin, rg->d_err, rg->wave_lds, rg->out_lds,
                                   rg->blocks, rg->waves, s);
    return e == hipSuccess ? FLS_OK : hip_fail(e, "row-group launch");
}

}  // namespace

extern "C" {

int fls_gpu_rowgroup_create(uint32_t n_columns, const uint8_t* const* host_columns, const uint64_t* column_sizes,
                            const uint64_t* device_offsets, const uint8_t* types, fls_gpu_rowgroup** out) {
    if (!out || (n_columns && (!host_columns || !column_sizes || !device_offsets || !types)))
        return fail(FLS_E_ARG, "null argument");
    *out    = nullptr;
    auto* rg = new fls_gpu_rowgroup();
    std::vector<const Column*> cols;
    for (uint32_t c = 0; c < n_columns; ++c) {
        if (types[c] == STR) return free_rg(rg), fail(FLS_E_UNSUPPORTED, "string columns use fls_gpu_plan_decode_strings");
        fls_gpu_plan* p  = nullptr;
        int           rc = fls_gpu_plan_create(host_columns[c], column_sizes[c], (enum data_t)types[c], &p);
        if (rc) return free_rg(rg), rc;
        rg->plans.push_back(p);
        rg->dev_off.push_back(device_offsets[c]);
        rg->types.push_back((uint8_t)types[c]);
        cols.push_back(&plan_column(p));
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    hipError_t      e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return free_rg(rg), hip_fail(e, "device properties");
    // at least 8 tiles per wave slot of the device (16 waves per compute unit)
    const uint64_t    min_tiles = 8ull * 16 * (uint64_t)prop.multiProcessorCount;
    WorkList          wl;
    const std::string bad = plan_work_list_sized(
        cols, host_columns, types, device_offsets, column_sizes, min_tiles,
        [&](size_t c, size_t i) { return (uint64_t)(uintptr_t)plan_dictionary(rg->plans[c], i); }, wl);
    if (!bad.empty()) return free_rg(rg), fail(FLS_E_FORMAT, "row-group work list: " + bad);
    rg->tiles     = std::move(wl.tiles);
    rg->fused     = std::move(wl.fused);
    rg->out_off   = std::move(wl.out_off);
    rg->out_total = wl.out_total;
    rg->wave_lds  = wl.wave_lds;
    // blocks of kRgWaves waves, each wave owning a run of the list: by default as many per compute unit as the
    // kernel's registers and LDS keep resident at once; the tuning table's max_blocks sets another count per compute
    // unit (more blocks than fit: shorter runs, the later blocks start as earlier ones end)
    // (a block holds as many of the kRgWaves waves as its LDS limit allows: each wave's slot is its staged ranges
    // plus an output buffer for one vector of the widest fused column)
    if (!rg->tiles.empty()) {
        for (const RgTile& t : rg->tiles) rg->out_lds = std::max(rg->out_lds, rg_out_lds(rg_kind_bits(t.kind)));
        const size_t slot = (size_t)rg->wave_lds + rg->out_lds;
        rg->waves         = (unsigned)std::min<size_t>(kRgWaves, prop.sharedMemPerBlock / slot);
        if (!rg->waves) return free_rg(rg), fail(FLS_E_UNSUPPORTED, "row-group tile exceeds the block's LDS");
        const Tune     tn     = tune_for(FAM_ROWGROUP, 0, 255);
        const unsigned fit    = rg_blocks_per_cu((unsigned)slot, rg->waves);
        const unsigned per_cu = std::max(1u, std::min(tn.max_blocks ? tn.max_blocks : fit, 4096u));
        rg->blocks            = (unsigned)prop.multiProcessorCount * per_cu;
        rg->wave_begin        = rg_assign_waves(rg->tiles, rg->blocks * rg->waves);
    }
    if ((e = hipMalloc(&rg->d_err, 4)) != hipSuccess) return free_rg(rg), hip_fail(e, "hipMalloc");
    if ((e = hipMemset(rg->d_err, 0, 4)) != hipSuccess) return free_rg(rg), hip_fail(e, "hipMemset");
    if (!rg->tiles.empty()) {
        const size_t b = rg->tiles.size() * sizeof(RgTile), wb = rg->wave_begin.size() * sizeof(uint32_t);
        if ((e = hipMalloc(&rg->d_tiles, b)) != hipSuccess) return free_rg(rg), hip_fail(e, "hipMalloc");
        if ((e = hipMemcpy(rg->d_tiles, rg->tiles.data(), b, hipMemcpyHostToDevice)) != hipSuccess)
            return free_rg(rg), hip_fail(e, "hipMemcpy");
        if ((e = hipMalloc(&rg->d_wave_begin, wb)) != hipSuccess) return free_rg(rg), hip_fail(e, "hipMalloc");
        if ((e = hipMemcpy(rg->d_wave_begin, rg->wave_begin.data(), wb, hipMemcpyHostToDevice)) != hipSuccess)
            return free_rg(rg), hip_fail(e, "hipMemcpy");
    }
    // parallel branches (default-priority, non-blocking streams) and the column order for them
    const unsigned nb = (unsigned)std::min<size_t>(4, std::max<size_t>(1, rg->plans.size()));
    if ((e = hipEventCreateWithFlags(&rg->fork, hipEventDisableTiming)) != hipSuccess) return free_rg(rg), hip_fail(e, "event");
    for (unsigned b = 0; b < nb; ++b) {
        hipStream_t st;
        hipEvent_t  ev;
        if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return free_rg(rg), hip_fail(e, "stream");
        rg->branch.push_back(st);
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return free_rg(rg), hip_fail(e, "event");
        rg->join.push_back(ev);
    }
    for (size_t c = 0; c < rg->plans.size(); ++c) rg->order.push_back(c);
    std::stable_sort(rg->order.begin(), rg->order.end(), [&](size_t a, size_t b) {
        auto work = [&](size_t c) {
            return fls_gpu_plan_n_values(rg->plans[c]) * (dtype_bits(rg->types[c]) / 8) + fls_gpu_plan_encoded_size(rg->plans[c]);
        };
        return work(a) > work(b);
    });
    *out                   = rg;
    return FLS_OK;
}

void fls_gpu_rowgroup_destroy(fls_gpu_rowgroup* rg) {
    if (rg) free_rg(rg);
}

int fls_gpu_rowgroup_layout(const fls_gpu_rowgroup* rg, uint64_t* out_offsets, uint64_t* out_total_bytes) {
    if (!rg) return fail(FLS_E_ARG, "null argument");
    if (out_offsets)
        for (size_t c = 0; c < rg->out_off.size(); ++c) out_offsets[c] = rg->out_off[c];
    if (out_total_bytes) *out_total_bytes = rg->out_total;
    return FLS_OK;
}

int fls_gpu_rowgroup_decode(fls_gpu_rowgroup* rg, const uint8_t* dev_encoded, uint8_t* dev_out, int mode, void* stream) {
    if (!rg || !dev_encoded || !dev_out) return fail(FLS_E_ARG, "null argument");
    if (((uintptr_t)dev_out & 255) != 0) return fail(FLS_E_ARG, "row-group output must be 256-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    if (mode == 0) return decode_fused(rg, dev_encoded, dev_out, s);
    if (mode == 1) return decode_pages(rg, dev_encoded, dev_out, s);
    if (mode == 2) return decode_pages_branched(rg, dev_encoded, dev_out, s);
    return fail(FLS_E_ARG, "mode must be 0 (persistent), 1 (per page) or 2 (per page, parallel branches)");
}

int fls_gpu_rowgroup_decode_graph(fls_gpu_rowgroup* rg, const uint8_t* dev_encoded, uint8_t* dev_out, void* stream) {
    if (!rg || !dev_encoded || !dev_out) return fail(FLS_E_ARG, "null argument");
    if (((uintptr_t)dev_out & 255) != 0) return fail(FLS_E_ARG, "row-group output must be 256-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    hipError_t  e;
    if (!rg->exec || rg->g_enc != dev_encoded || rg->g_out != dev_out) {
        if (rg->exec) (void)hipGraphExecDestroy(rg->exec), rg->exec = nullptr;
        if (rg->graph) (void)hipGraphDestroy(rg->graph), rg->graph = nullptr;
        hipStream_t cs;
        if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess) return hip_fail(e, "stream");
        if ((e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
            (void)hipStreamDestroy(cs);
            return hip_fail(e, "begin capture");
        }
        int rc = decode_fused(rg, dev_encoded, dev_out, cs);
        e      = hipStreamEndCapture(cs, &rg->graph);
        (void)hipStreamDestroy(cs);
        if (rc) return rc;
        if (e != hipSuccess) return hip_fail(e, "end capture");
        if ((e = hipGraphInstantiate(&rg->exec, rg->graph, nullptr, nullptr, 0)) != hipSuccess)
            return hip_fail(e, "graph instantiate");
        rg->g_enc = dev_encoded, rg->g_out = dev_out;
    }
    e = hipGraphLaunch(rg->exec, s);
    return e == hipSuccess ? FLS_OK : hip_fail(e, "graph launch");
}

int fls_rowgroup_plan_work_list(uint32_t n_columns, const uint8_t* const* host_columns, const uint64_t* column_sizes,
                                const uint64_t* device_offsets, const uint8_t* types, uint32_t waves, uint32_t* n_tiles,
                                void* tiles, uint32_t capacity, uint32_t* wave_begin, uint32_t* wave_lds) {
    if (n_columns && (!host_columns || !column_sizes || !device_offsets || !types)) return fail(FLS_E_ARG, "null argument");
    std::vector<Column>        parsed(n_columns);
    std::vector<const Column*> cols;
    for (uint32_t c = 0; c < n_columns; ++c) {
        if (types[c] == STR) return fail(FLS_E_UNSUPPORTED, "string columns use fls_gpu_plan_decode_strings");
        const std::string err = parse_column(host_columns[c], column_sizes[c], types[c], parsed[c]);
        if (!err.empty()) return fail(FLS_E_FORMAT, err);
        cols.push_back(&parsed[c]);
    }
    // no device dictionaries exist here: dictionary tiles name the page's dictionary by a nonzero placeholder
    WorkList          wl;
    const std::string bad = plan_work_list_sized(cols, host_columns, types, device_offsets, column_sizes, 8ull * 16 * 256,
                                                 [&](size_t c, size_t i) { return (uint64_t)(c * 65536 + i + 1); }, wl);
    if (!bad.empty()) return fail(FLS_E_FORMAT, "row-group work list: " + bad);
    if (waves) {
        const std::vector<uint32_t> begin = rg_assign_waves(wl.tiles, waves);
        if (wave_begin) std::memcpy(wave_begin, begin.data(), begin.size() * sizeof(uint32_t));
    }
    if (n_tiles) *n_tiles = (uint32_t)wl.tiles.size();
    if (tiles) std::memcpy(tiles, wl.tiles.data(), std::min<size_t>(capacity, wl.tiles.size()) * sizeof(RgTile));
    if (wave_lds) *wave_lds = wl.wave_lds;
    return FLS_OK;
}

int fls_gpu_rowgroup_work_list(const fls_gpu_rowgroup* rg, uint32_t* n_tiles, void* tiles, uint32_t capacity,
                               uint32_t* blocks, uint32_t* waves, uint32_t* wave_lds, uint32_t* wave_begin) {
    if (!rg) return fail(FLS_E_ARG, "null argument");
    if (wave_begin && !rg->wave_begin.empty())
        std::memcpy(wave_begin, rg->wave_begin.data(), rg->wave_begin.size() * sizeof(uint32_t));
    if (n_tiles) *n_tiles = (uint32_t)rg->tiles.size();
    if (tiles) std::memcpy(tiles, rg->tiles.data(), std::min<size_t>(capacity, rg->tiles.size()) * sizeof(RgTile));
    if (blocks) *blocks = rg->blocks;
    if (waves) *waves = rg->waves;
    if (wave_lds) *wave_lds = rg->wave_lds;
    return FLS_OK;
}

int fls_gpu_rowgroup_check(fls_gpu_rowgroup* rg, void* stream) {
    if (!rg) return fail(FLS_E_ARG, "null argument");
    uint32_t   h = 0;
    hipError_t e = hipMemcpyAsync(&h, rg->d_err, 4, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "check");
    if (h) return fail(FLS_E_DATA, "row-group decode met an out-of-range run or dictionary entry");
    for (auto* p : rg->plans) {
        int rc = fls_gpu_plan_check(p, stream);
        if (rc) return rc;
    }
    return FLS_OK;
}

}  // extern "C"
