// rowgroup.hip -- row-group decode: many columns, one persistent launch over a device-side work list.
//
// Replaces the per-column decode loop that sits behind Reader::get_chunk / Interpreter::Decoding in the reference
// (src/reader/reader.cpp:19-61, src/expression/interpreter.cpp:105-137) with a work list of (page, vector range)
// tiles decoded by one persistent kernel (k_rowgroup.hip); mode 1 keeps the one-launch-per-page form for
// comparison. The persistent launch can be captured in a HIP graph and replayed.
#include <algorithm>
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
    std::vector<RgTile>        tiles;
    RgTile*                    d_tiles   = nullptr;
    uint32_t*                  d_counter = nullptr;
    uint32_t*                  d_err     = nullptr;
    unsigned                   wave_lds = 0, blocks = 0;
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
                              uint64_t enc_size, uint64_t out_size) {
    // buffer count of each kind family: int UNC 1 / FFOR 3 / DELTA 4, RLE 5, dictionary codes 1 / 3 / 4
    static const unsigned kBuffers[7] = {1, 3, 4, 5, 1, 3, 4};
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
        const uint64_t vb = 128ull * rg_kind_bits(t.kind);  // bytes of one decoded vector
        if (t.out != pe->out || (uint64_t)t.vec_c * vb != pe->out_bytes) return why("output range differs from its page");
        if (t.out + (uint64_t)(t.v0 + t.nv) * vb > out_size) return why("output past the row-group's output");
        if (t.kind >= RG_DICT1_8 && (!t.dict || !t.dict_n)) return why("dictionary kind without a dictionary");
    }
    return "";
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
    case EXP_DELTA: return RG_DELTA8 + wc;
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
    if (rg->d_counter) (void)hipFree(rg->d_counter);
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

// The persistent launch's work list for parsed columns: tiles of about `tile_kib` KiB of decoded values per page
// of every fused column (a column is fused when all its pages have a kind), the LDS bytes per wave the launch
// needs, and the columns' output offsets. dict_of(c, i): device address of column c's page-i dictionary. The list
// is checked (rg_validate_tiles) before it is returned.
struct WorkList {
    std::vector<RgTile>   tiles;
    std::vector<bool>     fused;
    std::vector<uint64_t> out_off;
    uint64_t              out_total = 0;
    unsigned              wave_lds  = 0;
};
template <typename DictOf>
std::string plan_work_list(const std::vector<const Column*>& cols, const uint8_t* types, const uint64_t* device_offsets,
                           const uint64_t* column_sizes, unsigned tile_kib, DictOf&& dict_of, WorkList& wl) {
    std::vector<RgPageExtent> extents;
    uint64_t                  enc_size = 0;
    for (size_t c = 0; c < cols.size(); ++c) {
        const Column&  col   = *cols[c];
        const unsigned W     = dtype_bits(types[c]);
        const uint64_t bytes = col.n_values * (W / 8);
        enc_size             = std::max(enc_size, device_offsets[c] + column_sizes[c]);
        wl.out_off.push_back(wl.out_total);
        bool fused = true;
        for (const auto& pe : col.pages) fused &= kind_of(pe, W) >= 0;
        wl.fused.push_back(fused);
        if (fused) {
            const unsigned tile_vec =
                std::max(1u, std::min(kRgMaxTileVec, (unsigned)((tile_kib * 1024ull) / (1024ull * (W / 8)))));
            for (size_t i = 0; i < col.pages.size(); ++i) {
                const PageEntry& pe   = col.pages[i];
                const int        kind = kind_of(pe, W);
                wl.wave_lds           = std::max(wl.wave_lds, rg_wave_lds(kind));
                extents.push_back({device_offsets[c] + pe.page_off, pe.page_size, wl.out_total + pe.out_index * (W / 8),
                                   (uint64_t)pe.vec_c * 1024 * (W / 8)});
                for (uint32_t v0 = 0; v0 < pe.vec_c; v0 += tile_vec) {
                    RgTile t {};
                    t.page   = device_offsets[c] + pe.page_off;
                    t.out    = wl.out_total + pe.out_index * (W / 8);
                    t.dict   = dict_of(c, i);
                    t.dict_n = pe.dict_n;
                    t.v0     = v0;
                    t.nv     = (uint16_t)std::min(tile_vec, pe.vec_c - v0);
                    t.vec_c  = pe.vec_c;
                    t.kind   = (uint16_t)kind;
                    t.arr_c  = (uint16_t)pe.arr_c;
                    for (unsigned k = 0; k < 5 && k < pe.arr_c; ++k) t.off[k] = pe.boff[k];
                    wl.tiles.push_back(t);
                }
            }
        }
        wl.out_total += (bytes + 255) & ~255ull;
    }
    // widest values first, each kind's tiles together: the long tiles start early and the short 8-bit ones fill
    // the tail, and a compute unit's consecutive tiles mostly share one code path
    std::stable_sort(wl.tiles.begin(), wl.tiles.end(), [](const RgTile& a, const RgTile& b) {
        const unsigned wa = rg_kind_bits(a.kind), wb = rg_kind_bits(b.kind);
        return wa != wb ? wa > wb : a.kind < b.kind;
    });
    std::string bad = rg_validate_tiles(wl.tiles, extents, enc_size, wl.out_total);
    if (!bad.empty()) return bad;
    for (const RgTile& t : wl.tiles)
        if (rg_wave_lds(t.kind) > wl.wave_lds) return "LDS size";
    return "";
}

unsigned tile_kib_tuned() {
    const Tune tn = tune_for(FAM_ROWGROUP, 0, 255);  // block = KiB of decoded values per tile
    return tn.block ? tn.block : 64;
}

// Plans with the tuned tile size, then with smaller tiles while the list has fewer than `min_tiles` tiles (a small
// row-group would otherwise leave most waves without work), down to 16 KiB per tile.
template <typename DictOf>
std::string plan_work_list_sized(const std::vector<const Column*>& cols, const uint8_t* types,
                                 const uint64_t* device_offsets, const uint64_t* column_sizes, uint64_t min_tiles,
                                 DictOf&& dict_of, WorkList& wl) {
    unsigned tile_kib = tile_kib_tuned();
    for (;;) {
        wl = WorkList {};
        std::string bad = plan_work_list(cols, types, device_offsets, column_sizes, tile_kib, dict_of, wl);
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
    hipError_t e = launch_rowgroup(enc, out, rg->d_tiles, (uint32_t)rg->tiles.size(), rg->d_counter, rg->d_err,
                                   rg->wave_lds, rg->blocks, s);
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
        cols, types, device_offsets, column_sizes, min_tiles,
        [&](size_t c, size_t i) { return (uint64_t)(uintptr_t)plan_dictionary(rg->plans[c], i); }, wl);
    if (!bad.empty()) return free_rg(rg), fail(FLS_E_FORMAT, "row-group work list: " + bad);
    rg->tiles     = std::move(wl.tiles);
    rg->fused     = std::move(wl.fused);
    rg->out_off   = std::move(wl.out_off);
    rg->out_total = wl.out_total;
    rg->wave_lds  = wl.wave_lds;
    if ((e = hipMalloc(&rg->d_counter, 4)) != hipSuccess) return free_rg(rg), hip_fail(e, "hipMalloc");
    if ((e = hipMalloc(&rg->d_err, 4)) != hipSuccess) return free_rg(rg), hip_fail(e, "hipMalloc");
    if ((e = hipMemset(rg->d_err, 0, 4)) != hipSuccess) return free_rg(rg), hip_fail(e, "hipMemset");
    if (!rg->tiles.empty()) {
        const size_t b = rg->tiles.size() * sizeof(RgTile);
        if ((e = hipMalloc(&rg->d_tiles, b)) != hipSuccess) return free_rg(rg), hip_fail(e, "hipMalloc");
        if ((e = hipMemcpy(rg->d_tiles, rg->tiles.data(), b, hipMemcpyHostToDevice)) != hipSuccess)
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
    // blocks of kRgWaves waves: as many per compute unit as the LDS allows (tuning max_blocks caps it)
    const unsigned lds_cu    = prop.maxSharedMemoryPerMultiProcessor ? (unsigned)prop.maxSharedMemoryPerMultiProcessor : 65536u;
    const unsigned per_block = rg->wave_lds * kRgWaves;
    const unsigned fit       = per_block ? std::max(1u, lds_cu / per_block) : 8u;
    const Tune     tn        = tune_for(FAM_ROWGROUP, 0, 255);
    const unsigned per_cu    = std::min(tn.max_blocks ? tn.max_blocks : 8u, fit);
    rg->blocks               = (unsigned)prop.multiProcessorCount * per_cu;
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
                                const uint64_t* device_offsets, const uint8_t* types, uint32_t* n_tiles, void* tiles,
                                uint32_t capacity, uint32_t* wave_lds) {
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
    const std::string bad = plan_work_list_sized(cols, types, device_offsets, column_sizes, 8ull * 16 * 256,
                                                 [&](size_t c, size_t i) { return (uint64_t)(c * 65536 + i + 1); }, wl);
    if (!bad.empty()) return fail(FLS_E_FORMAT, "row-group work list: " + bad);
    if (n_tiles) *n_tiles = (uint32_t)wl.tiles.size();
    if (tiles) std::memcpy(tiles, wl.tiles.data(), std::min<size_t>(capacity, wl.tiles.size()) * sizeof(RgTile));
    if (wave_lds) *wave_lds = wl.wave_lds;
    return FLS_OK;
}

int fls_gpu_rowgroup_work_list(const fls_gpu_rowgroup* rg, uint32_t* n_tiles, void* tiles, uint32_t capacity,
                               uint32_t* blocks, uint32_t* wave_lds) {
    if (!rg) return fail(FLS_E_ARG, "null argument");
    if (n_tiles) *n_tiles = (uint32_t)rg->tiles.size();
    if (tiles) std::memcpy(tiles, rg->tiles.data(), std::min<size_t>(capacity, rg->tiles.size()) * sizeof(RgTile));
    if (blocks) *blocks = rg->blocks;
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
