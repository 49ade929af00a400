// gpu_api.hip -- the C-ABI of include/fls_gpu.h: column plans, launches, error reporting.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/fls_gpu.h"
#include "format.hpp"
#include "launch.hpp"
#include "page_kernels.hpp"
#include "gpu_internal.hpp"

namespace fls {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
int hip_fail(hipError_t e, const char* what) {
    return fail(FLS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace fls

using namespace fls;

struct DevDict {
    uint8_t*  values = nullptr;  // decoded dictionary (DW-bit values)
    HzlChunk* chunks = nullptr;
    uint32_t  n_chunks = 0;
    bool      exc      = false;  // some chunk carries exceptions
    // DICT_RLE: the page's run codes (HZL chunks), decoded codes and their dictionary values
    HzlChunk* run_chunks   = nullptr;
    uint32_t  n_run_chunks = 0;
    bool      run_exc      = false;
    uint32_t* codes        = nullptr;
    // DICT_RLE: run values; other forms with a chain in buffer 0 (S_PATCH_REC exceptions, RLE run values): its
    // decoded bytes
    uint8_t* run_values = nullptr;
    // chain in buffer 1 (B_PATCH_REC exceptions, RLE position lane bases): chunks and decoded bytes
    HzlChunk* rec1_chunks   = nullptr;
    uint32_t  n_rec1_chunks = 0;
    bool      rec1_exc      = false;
    uint8_t*  rec1_values   = nullptr;
};

namespace {
// Device chunk table of an HZL chain of dw-bit values. The raw tail is split into pieces of 1024 values, so that a
// long tail (a chain whose values need the full width) is copied by many blocks; every entry keeps dst = 1024 * index.
std::vector<HzlChunk> device_chunks(const std::vector<HzlChunkH>& h, bool& exc, unsigned dw) {
    std::vector<HzlChunk> ch;
    ch.reserve(h.size());
    for (const auto& c : h) {
        exc |= c.exc_c != 0;
        if (c.bw != 255) {
            ch.push_back({c.src, c.dst, c.base, c.bw, c.exc_c, c.n, 0});
            continue;
        }
        uint32_t k = 0;
        do {
            const uint32_t n = c.n - k < 1024 ? c.n - k : 1024;
            ch.push_back({c.src + (uint64_t)k * (dw / 8), c.dst + k, c.base, 255u, 0u, n, 0});
            k += n;
        } while (k < c.n);
    }
    return ch;
}
}  // namespace

struct fls_gpu_plan {
    Column               col;
    unsigned             W  = 0;
    unsigned             DW = 0;
    uint32_t*            d_err = nullptr;
    std::vector<DevDict> dicts;  // one per page entry (empty for non-dictionary pages)
    StringPlan           strings;
};

namespace fls {
const Column& plan_column(const fls_gpu_plan* p) { return p->col; }
const void*   plan_dictionary(const fls_gpu_plan* p, size_t page) {
    return page < p->dicts.size() ? p->dicts[page].values : nullptr;
}
int plan_decode_dictionaries(fls_gpu_plan* p, const uint8_t* dev_encoded, hipStream_t s) {
    for (size_t i = 0; i < p->col.pages.size(); ++i) {
        const DevDict& d = p->dicts[i];
        if (!d.n_chunks) continue;
        hipError_t e = launch_hzl(p->DW, dev_encoded, d.chunks, d.n_chunks, d.values, d.exc, s);
        if (e != hipSuccess) return hip_fail(e, "dictionary launch");
    }
    return FLS_OK;
}
}  // namespace fls

namespace {
// DICT_RLE page: dictionary and run codes from their HZL chains, run values = dict[code], then the RLE kernel
// (positions DELTA_NO_PATCH) gathers from those values.
hipError_t drle_page(fls_gpu_plan* plan, size_t i, const uint8_t* dev_encoded, const uint8_t* page, const PageMeta& pm,
                     uint8_t* out, const SelArgs* sel, hipStream_t s) {
    const PageEntry& pe = plan->col.pages[i];
    DevDict&         d  = plan->dicts[i];
    if (!pe.vec_c) return hipSuccess;
    hipError_t e = hipSuccess;
    if (!d.exc && !d.run_exc) {  // one pass: codes and their values straight from the two chains
        e = launch_drle_fused(plan->W, dev_encoded, d.run_chunks, d.n_run_chunks, d.chunks, d.n_chunks, pe.dict_n,
                              d.run_values, plan->d_err, s);
    } else {
        e = launch_hzl(plan->DW, dev_encoded, d.chunks, d.n_chunks, d.values, d.exc, s);
        if (e == hipSuccess) e = launch_hzl(32, dev_encoded, d.run_chunks, d.n_run_chunks, (uint8_t*)d.codes, d.run_exc, s);
        if (e == hipSuccess)
            e = launch_drle_values(plan->W, d.codes, pe.run_codes, d.values, pe.dict_n, d.run_values, plan->d_err, s);
    }
    const RunSrc rs {d.run_values, (uint32_t)(4 * pe.run_codes), 2u};
    if (e == hipSuccess) e = launch_page_rle(plan->W, 2, page, pm, pe.vec_c, out, plan->d_err, s, sel, &rs, pe.pos_bw_max <= 1);
    return e;
}

// Decoded bytes of the page's chained buffers (buffer 0, buffer 1), written by one HZL launch each.
hipError_t decode_chains(fls_gpu_plan* plan, size_t i, const uint8_t* dev_encoded, hipStream_t s) {
    const PageEntry& pe = plan->col.pages[i];
    DevDict&         d  = plan->dicts[i];
    hipError_t       e  = hipSuccess;
    if (d.n_run_chunks && pe.exp_t != EXP_DICT_RLE)
        e = launch_hzl(pe.rec_width, dev_encoded, d.run_chunks, d.n_run_chunks, d.run_values, d.run_exc, s);
    if (e == hipSuccess && d.n_rec1_chunks)
        e = launch_hzl(pe.rec_width, dev_encoded, d.rec1_chunks, d.n_rec1_chunks, d.rec1_values, d.rec1_exc, s);
    return e;
}

// One page: decode into `out`, or (sel) write its selection bitmaps.
hipError_t page_launch(fls_gpu_plan* plan, size_t i, const uint8_t* dev_encoded, uint8_t* out, const SelArgs* sel,
                       hipStream_t s) {
    const PageEntry& pe   = plan->col.pages[i];
    const unsigned   W    = plan->W;
    const uint8_t*   page = dev_encoded + pe.page_off;
    DevDict&         d    = plan->dicts[i];
    PageMeta         pm {};
    for (unsigned b = 0; b < pe.arr_c && b < 6; ++b) pm.off[b] = pe.boff[b];
    pm.arr_c = pe.arr_c, pm.fixed = pe.fixed_meta;
    if (!pe.vec_c) return hipSuccess;
    Expr e;
    expr_from_page(pe.exp_t, pe.var, e);
    hipError_t     r     = pe.rec ? decode_chains(plan, i, dev_encoded, s) : hipSuccess;
    const uint8_t  patch = expr_patch(e);
    const bool     rsum  = expr_rsum(e);
    // exceptions of a *_REC form: the decoded chain of their buffer (0 without lane bases, 1 with them)
    const uint8_t* exc = (pe.rec & 1) ? d.run_values : (pe.rec & 2) ? d.rec1_values : nullptr;
    if (r != hipSuccess) return r;
    switch (pe.exp_t) {
    case EXP_UNCOMPRESSED: return launch_page_int(W, 0, page, pm, pe.vec_c, out, s, sel);
    case EXP_FFOR:
    case EXP_DELTA:
        if (!patch) return launch_page_int(W, rsum ? 2 : 1, page, pm, pe.vec_c, out, s, sel);
        return launch_page_patch(W, patch, rsum, nullptr, 0, page, pm, pe.vec_c, out, plan->d_err, s, sel, exc);
    case EXP_RLE: {
        const int kv = pe.var == 1 || pe.var == kRleRec ? 1 : 2;  // positions: plain u16, or RSUM + FFOR
        const bool bw1 = pe.pos_bw_max <= 1;
        if (!pe.rec) return launch_page_rle(W, kv, page, pm, pe.vec_c, out, plan->d_err, s, sel, nullptr, bw1);
        const RunSrc rs {d.run_values, (uint32_t)(8 * pe.run_codes), 3u, (pe.rec & 2) ? d.rec1_values : nullptr};
        return launch_page_rle(W, kv, page, pm, pe.vec_c, out, plan->d_err, s, sel, &rs, bw1);
    }
    case EXP_DICT_RLE: return drle_page(plan, i, dev_encoded, page, pm, out, sel, s);
    default: {
        const HzlSrc hz {dev_encoded, d.chunks, d.n_chunks, d.exc ? 1u : 0u};
        if (pe.var == 1 || pe.var == 2 || pe.var == kDictDelta)
            return launch_page_dict(W, pe.var, page, pm, pe.vec_c, d.values, pe.dict_n, out, plan->d_err, s, sel, &hz);
        // patched code pages: the dictionary is decoded first, then the patch kernel gathers from it
        r = launch_hzl(plan->DW, dev_encoded, d.chunks, d.n_chunks, d.values, d.exc, s);
        if (r != hipSuccess) return r;
        return launch_page_patch(W, patch, rsum, d.values, pe.dict_n, page, pm, pe.vec_c, out, plan->d_err, s, sel,
                                 exc);
    }
    }
}
}  // namespace

extern "C" {

const char* fls_gpu_error_string(void) { return g_err.c_str(); }

int fls_gpu_plan_create(const uint8_t* host_encoded, uint64_t encoded_size, enum data_t type,
                        fls_gpu_plan** out_plan) {
    if (!host_encoded || !out_plan) return fail(FLS_E_ARG, "null argument");
    *out_plan = nullptr;
    auto plan = std::make_unique<fls_gpu_plan>();
    auto err  = parse_column(host_encoded, encoded_size, (uint8_t)type, plan->col);
    if (!err.empty()) return fail(FLS_E_FORMAT, err);
    plan->W  = type == STR ? 8 : dtype_bits((uint8_t)type);
    plan->DW = plan->W < 32 ? 32 : plan->W;
    hipError_t e;
    if ((e = hipMalloc(&plan->d_err, 4)) != hipSuccess) return hip_fail(e, "hipMalloc");
    if ((e = hipMemset(plan->d_err, 0, 4)) != hipSuccess) return hip_fail(e, "hipMemset");
    plan->dicts.resize(plan->col.pages.size());
    for (size_t i = 0; i < plan->col.pages.size(); ++i) {
        const PageEntry& pe = plan->col.pages[i];
        DevDict&         d  = plan->dicts[i];
        if (!pe.hzl.empty()) {
            std::vector<HzlChunk> ch = device_chunks(pe.hzl, d.exc, plan->DW);
            d.n_chunks               = (uint32_t)ch.size();
            const uint64_t vals = (uint64_t)pe.dict_n + 1024;  // chunk writes are whole 1024-value vectors
            if ((e = hipMalloc(&d.values, vals * (plan->DW / 8))) != hipSuccess) return hip_fail(e, "hipMalloc");
            if ((e = hipMalloc(&d.chunks, ch.size() * sizeof(HzlChunk))) != hipSuccess) return hip_fail(e, "hipMalloc");
            if ((e = hipMemcpy(d.chunks, ch.data(), ch.size() * sizeof(HzlChunk), hipMemcpyHostToDevice)) != hipSuccess)
                return hip_fail(e, "hipMemcpy");
        }
        if (!pe.run_hzl.empty()) {  // chain in buffer 0: DICT_RLE run codes, S_PATCH_REC exceptions, RLE run values
            std::vector<HzlChunk> rc = device_chunks(pe.run_hzl, d.run_exc, pe.rec_width);
            d.n_run_chunks           = (uint32_t)rc.size();
            if ((e = hipMalloc(&d.run_chunks, rc.size() * sizeof(HzlChunk))) != hipSuccess) return hip_fail(e, "hipMalloc");
            if ((e = hipMemcpy(d.run_chunks, rc.data(), rc.size() * sizeof(HzlChunk), hipMemcpyHostToDevice)) != hipSuccess)
                return hip_fail(e, "hipMemcpy");
            if (pe.exp_t == EXP_DICT_RLE) {
                // chunk writes are whole 1024-code vectors
                if ((e = hipMalloc(&d.codes, (pe.run_codes + 1024) * 4)) != hipSuccess) return hip_fail(e, "hipMalloc");
                if ((e = hipMalloc(&d.run_values, pe.run_codes * (plan->W / 8) + 16)) != hipSuccess)
                    return hip_fail(e, "hipMalloc");
            } else if ((e = hipMalloc(&d.run_values, (pe.run_codes + 1024) * (pe.rec_width / 8))) != hipSuccess) {
                return hip_fail(e, "hipMalloc");  // decoded chain bytes
            }
        }
        if (!pe.rec1_hzl.empty()) {  // chain in buffer 1: B_PATCH_REC exceptions, RLE position lane bases
            std::vector<HzlChunk> rc = device_chunks(pe.rec1_hzl, d.rec1_exc, pe.rec_width);
            d.n_rec1_chunks          = (uint32_t)rc.size();
            if ((e = hipMalloc(&d.rec1_chunks, rc.size() * sizeof(HzlChunk))) != hipSuccess) return hip_fail(e, "hipMalloc");
            if ((e = hipMemcpy(d.rec1_chunks, rc.data(), rc.size() * sizeof(HzlChunk), hipMemcpyHostToDevice)) != hipSuccess)
                return hip_fail(e, "hipMemcpy");
            if ((e = hipMalloc(&d.rec1_values, (pe.rec1_n + 1024) * (pe.rec_width / 8))) != hipSuccess)
                return hip_fail(e, "hipMalloc");
        }
    }
    if (type == STR) {
        int rc = strings_prepare(plan->col, host_encoded, plan->strings);
        if (rc) return rc;
    }
    *out_plan = plan.release();
    return FLS_OK;
}

void fls_gpu_plan_destroy(fls_gpu_plan* plan) {
    if (!plan) return;
    for (auto& d : plan->dicts) {
        if (d.values) (void)hipFree(d.values);
        if (d.chunks) (void)hipFree(d.chunks);
        if (d.run_chunks) (void)hipFree(d.run_chunks);
        if (d.codes) (void)hipFree(d.codes);
        if (d.run_values) (void)hipFree(d.run_values);
        if (d.rec1_chunks) (void)hipFree(d.rec1_chunks);
        if (d.rec1_values) (void)hipFree(d.rec1_values);
    }
    strings_release(plan->strings);
    if (plan->d_err) (void)hipFree(plan->d_err);
    delete plan;
}

uint64_t fls_gpu_plan_n_values(const fls_gpu_plan* plan) { return plan ? plan->col.n_values : 0; }
uint64_t fls_gpu_plan_encoded_size(const fls_gpu_plan* plan) { return plan ? plan->col.size : 0; }
uint32_t fls_gpu_plan_n_pages(const fls_gpu_plan* plan) { return plan ? (uint32_t)plan->col.pages.size() : 0; }

int fls_gpu_plan_decode(fls_gpu_plan* plan, const uint8_t* dev_encoded, void* dev_out, uint64_t out_capacity,
                        void* stream) {
    if (!plan || !dev_encoded || (!dev_out && plan->col.n_values)) return fail(FLS_E_ARG, "null argument");
    if (plan->col.dtype == STR) return fail(FLS_E_ARG, "string column: use fls_gpu_plan_decode_strings");
    if (((uintptr_t)dev_out & 15) != 0) return fail(FLS_E_ARG, "output must be 16-byte aligned");
    const unsigned W = plan->W;
    if (out_capacity < plan->col.n_values * (W / 8)) return fail(FLS_E_CAPACITY, "output capacity too small");
    hipStream_t s = (hipStream_t)stream;
    for (size_t i = 0; i < plan->col.pages.size(); ++i) {
        const PageEntry& pe  = plan->col.pages[i];
        uint8_t*         out = (uint8_t*)dev_out + pe.out_index * (W / 8);
        const hipError_t e   = page_launch(plan, i, dev_encoded, out, nullptr, s);
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
    }
    return FLS_OK;
}

int fls_gpu_plan_select(fls_gpu_plan* plan, const uint8_t* dev_encoded, int op, uint64_t constant,
                        uint64_t* dev_bitmap, uint32_t* dev_counts, void* stream) {
    if (!plan || !dev_encoded || ((!dev_bitmap || !dev_counts) && plan->col.n_values)) return fail(FLS_E_ARG, "null argument");
    if (plan->col.dtype == STR) return fail(FLS_E_UNSUPPORTED, "predicates on string columns are not supported");
    if (op < FLS_PRED_EQ || op > FLS_PRED_NE) return fail(FLS_E_ARG, "unknown predicate operator");
    const unsigned W = plan->W;
    const uint8_t  dt = plan->col.dtype;
    SelArgs        sa {};
    // one range test on an order-preserving key (see SelArgs)
    const uint64_t maxk = W == 64 ? ~0ull : ((1ull << W) - 1);
    uint64_t       key  = constant & maxk;
    if (dt == DT_DOUBLE) {
        sa.dbl     = 1;
        sa.nan_val = op == FLS_PRED_NE;
        const uint64_t mag = constant & 0x7fffffffffffffffull;
        if (mag > 0x7ff0000000000000ull) {  // NaN constant: every comparison false, != true
            sa.none = 1, sa.none_val = op == FLS_PRED_NE;
        }
        const uint64_t z = mag == 0 ? 0ull : constant;
        key              = (z >> 63) ? ~z : (z | 0x8000000000000000ull);
    } else if (dt == DT_I8 || dt == DT_I16 || dt == DT_I32 || dt == DT_I64) {
        sa.flip = 1ull << (W - 1);
        key ^= sa.flip;
    }
    switch (op) {
    case FLS_PRED_EQ: sa.lo = key, sa.span = 0; break;
    case FLS_PRED_NE: sa.lo = key, sa.span = 0, sa.inv = 1; break;
    case FLS_PRED_GT:
        if (key == maxk) sa.none = 1;
        else sa.lo = key + 1, sa.span = maxk - key - 1;
        break;
    case FLS_PRED_LT:
        if (key == 0) sa.none = 1;
        else sa.lo = 0, sa.span = key - 1;
        break;
    case FLS_PRED_GE: sa.lo = key, sa.span = maxk - key; break;
    default: sa.lo = 0, sa.span = key; break;  // FLS_PRED_LE
    }
    hipStream_t s = (hipStream_t)stream;
    for (size_t i = 0; i < plan->col.pages.size(); ++i) {
        const PageEntry& pe = plan->col.pages[i];
        const uint64_t   v0 = pe.out_index / 1024;
        SelArgs          ps = sa;
        ps.bitmap = dev_bitmap + v0 * 16, ps.counts = dev_counts + v0;
        const hipError_t e = page_launch(plan, i, dev_encoded, nullptr, &ps, s);
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
    }
    return FLS_OK;
}

int fls_gpu_plan_check(fls_gpu_plan* plan, void* stream) {
    if (!plan) return fail(FLS_E_ARG, "null plan");
    if (plan->col.dtype == STR) return strings_check(plan->strings, (hipStream_t)stream);
    uint32_t   h = 0;
    hipError_t e = hipMemcpyAsync(&h, plan->d_err, 4, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "check");
    if (h) return fail(FLS_E_DATA, "decoded data referenced an out-of-range run, dictionary entry or exception position");
    return FLS_OK;
}

int fls_gpu_plan_string_offsets(fls_gpu_plan* plan, const uint8_t* dev_encoded, uint64_t* dev_offsets,
                                void* stream) {
    if (!plan || !dev_encoded || !dev_offsets) return fail(FLS_E_ARG, "null argument");
    if (plan->col.dtype != STR) return fail(FLS_E_ARG, "not a string column");
    return strings_offsets(plan->col, plan->strings, dev_encoded, dev_offsets, (hipStream_t)stream);
}

int fls_gpu_plan_decode_strings(fls_gpu_plan* plan, const uint8_t* dev_encoded, const uint64_t* dev_offsets,
                                uint8_t* dev_bytes, uint64_t bytes_capacity, void* stream) {
    if (!plan || !dev_encoded || !dev_offsets || !dev_bytes) return fail(FLS_E_ARG, "null argument");
    if (plan->col.dtype != STR) return fail(FLS_E_ARG, "not a string column");
    return strings_decode(plan->col, plan->strings, dev_encoded, dev_offsets, dev_bytes, bytes_capacity,
                          (hipStream_t)stream);
}

int fls_gpu_plan_decode_strings_fused(fls_gpu_plan* plan, const uint8_t* dev_encoded, uint64_t* dev_offsets,
                                      uint8_t* dev_bytes, uint64_t bytes_capacity, void* stream) {
    if (!plan || !dev_encoded || !dev_offsets || !dev_bytes) return fail(FLS_E_ARG, "null argument");
    if (plan->col.dtype != STR) return fail(FLS_E_ARG, "not a string column");
    return strings_decode_all(plan->col, plan->strings, dev_encoded, dev_offsets, dev_bytes, bytes_capacity,
                              (hipStream_t)stream);
}

int fls_encode_strings(const uint32_t* offsets, const uint8_t* bytes, uint64_t n, uint8_t* out, uint64_t capacity,
                       uint64_t* out_size) {
    if (!offsets || (!bytes && n) || !out || !out_size) return fail(FLS_E_ARG, "null argument");
    const int64_t r = encode_str_column(13, offsets, bytes, n, out, capacity);
    if (r == -2) return fail(FLS_E_CAPACITY, "encoded column exceeds capacity");
    if (r < 0) return fail(FLS_E_UNSUPPORTED, "string count must be a multiple of 1024");
    *out_size = (uint64_t)r;
    return FLS_OK;
}

int fls_gpu_tune_set(int family, unsigned width, unsigned bw, unsigned block, unsigned nontemporal,
                     unsigned max_blocks) {
    if (family < 0 || family >= FAM_COUNT) return fail(FLS_E_ARG, "unknown kernel family");
    if (family != FAM_ROWGROUP && (block == 0 || block % 64 || block > kMaxBlock))
        return fail(FLS_E_ARG, "block must be a multiple of 64, <= 512");
    if (family == FAM_ROWGROUP && block == 0) return fail(FLS_E_ARG, "tile size must be positive");
    (void)tune_for(family, width, bw);  // make sure the default table is loaded first
    tune_set(family, width, bw, Tune {(uint16_t)block, (uint16_t)(nontemporal & 7), max_blocks});
    return FLS_OK;
}

int fls_gpu_tune_get(int family, unsigned width, unsigned bw, unsigned* block, unsigned* nontemporal,
                     unsigned* max_blocks) {
    if (family < 0 || family >= FAM_COUNT) return fail(FLS_E_ARG, "unknown kernel family");
    const Tune t = tune_for(family, width, bw);
    if (block) *block = t.block;
    if (nontemporal) *nontemporal = t.nt;
    if (max_blocks) *max_blocks = t.max_blocks;
    return FLS_OK;
}

int fls_gpu_tune_load(const char* path) {
    if (!path) return fail(FLS_E_ARG, "null path");
    (void)tune_for(0, 8, 0);
    int n = tune_load(path);
    return n < 0 ? fail(FLS_E_ARG, "cannot read tuning file") : n;
}

int fls_gpu_unffor_stream(unsigned width, unsigned bw, const void* dev_packed, const void* dev_bases, void* dev_out,
                          uint64_t n_vec, void* stream) {
    if (width != 8 && width != 16 && width != 32 && width != 64) return fail(FLS_E_ARG, "width must be 8/16/32/64");
    if (bw > width) return fail(FLS_E_ARG, "bit width larger than the element width");
    if (n_vec && (!dev_out || (bw && !dev_packed))) return fail(FLS_E_ARG, "null argument");
    if (((uintptr_t)dev_out & 15) != 0) return fail(FLS_E_ARG, "output must be 16-byte aligned");
    hipError_t e = launch_unffor_stream(width, bw, (const uint8_t*)dev_packed, dev_bases, (uint8_t*)dev_out, n_vec,
                                        (hipStream_t)stream);
    return e == hipSuccess ? FLS_OK : hip_fail(e, "kernel launch");
}

}  // extern "C"
