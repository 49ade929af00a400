// k_page.hip -- width dispatch for page kernels, dictionary payload kernels and the tuning table.
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <mutex>
#include <tuple>

#include "launch.hpp"
#include "page_kernels.hpp"

namespace fls {

namespace {
std::mutex                                          g_tune_mu;
std::map<std::tuple<int, unsigned, unsigned>, Tune> g_tune;
std::once_flag                                      g_tune_once;

// The tuning table shipped next to the library (or named by FLS_TUNING) is loaded on first use.
void load_default_tuning() {
    const char* env = std::getenv("FLS_TUNING");
    if (env && *env) {
        tune_load(env);
        return;
    }
    Dl_info info;
    if (dladdr((void*)&load_default_tuning, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        p = p.substr(0, p.find_last_of('/') + 1) + "tuning.txt";
        tune_load(p.c_str());
    }
}
}  // namespace

Tune tune_for(int family, unsigned W, unsigned bw) {
    std::call_once(g_tune_once, load_default_tuning);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto                        it = g_tune.find({family, W, bw});
    if (it != g_tune.end()) return it->second;
    it = g_tune.find({family, W, 255u});
    if (it != g_tune.end()) return it->second;
    return Tune {};
}
void tune_set(int family, unsigned W, unsigned bw, Tune t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[{family, W, bw}] = t;
}
int tune_load(const char* path) {
    FILE* f = std::fopen(path, "r");
    if (!f) return -1;
    char line[256];
    int  n = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        int      fam;
        unsigned W, bw, block, nt, maxb;
        if (std::sscanf(line, "%d %u %u %u %u %u", &fam, &W, &bw, &block, &nt, &maxb) != 6) continue;
        if (fam < 0 || fam >= FAM_COUNT || block == 0) continue;
        if (fam != FAM_ROWGROUP && (block % 64 || block > kMaxBlock)) continue;
        tune_set(fam, W, bw, Tune {(uint16_t)block, (uint16_t)(nt & 7), maxb});  // bits: 0 non-temporal, 1 direct, 2 kernel choice
        ++n;
    }
    std::fclose(f);
    return n;
}

hipError_t launch_page_int(unsigned W, int mode, const uint8_t* page, const PageMeta& pm, uint32_t vec_c, uint8_t* out,
                           hipStream_t s, const SelArgs* sel) {
    if (!vec_c) return hipSuccess;
    switch (W) {
    case 8: return page_int_w<8>(mode, page, pm, vec_c, out, sel, s);
    case 16: return page_int_w<16>(mode, page, pm, vec_c, out, sel, s);
    case 32: return page_int_w<32>(mode, page, pm, vec_c, out, sel, s);
    case 64: return page_int_w<64>(mode, page, pm, vec_c, out, sel, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_page_patch(unsigned W, int var, bool rsum, const void* dict, uint32_t dict_n, const uint8_t* page,
                             const PageMeta& pm, uint32_t vec_c, uint8_t* out, uint32_t* err, hipStream_t s,
                             const SelArgs* sel, const uint8_t* exc) {
    if (!vec_c) return hipSuccess;
    switch (W) {
    case 8: return page_patch_w<8>(var, rsum, dict, dict_n, page, pm, vec_c, out, err, sel, exc, s);
    case 16: return page_patch_w<16>(var, rsum, dict, dict_n, page, pm, vec_c, out, err, sel, exc, s);
    case 32: return page_patch_w<32>(var, rsum, dict, dict_n, page, pm, vec_c, out, err, sel, exc, s);
    case 64: return page_patch_w<64>(var, rsum, dict, dict_n, page, pm, vec_c, out, err, sel, exc, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_page_rle(unsigned W, int var, const uint8_t* page, const PageMeta& pm, uint32_t vec_c, uint8_t* out,
                           uint32_t* err, hipStream_t s, const SelArgs* sel, const RunSrc* runs, bool pos_bw1) {
    if (!vec_c) return hipSuccess;
    switch (W) {
    case 8: return page_rle_w<8>(var, page, pm, vec_c, out, err, sel, runs, pos_bw1, s);
    case 16: return page_rle_w<16>(var, page, pm, vec_c, out, err, sel, runs, pos_bw1, s);
    case 32: return page_rle_w<32>(var, page, pm, vec_c, out, err, sel, runs, pos_bw1, s);
    case 64: return page_rle_w<64>(var, page, pm, vec_c, out, err, sel, runs, pos_bw1, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_drle_values(unsigned W, const uint32_t* codes, uint64_t n, const void* dict, uint32_t dict_n,
                              uint8_t* rv, uint32_t* err, hipStream_t s) {
    if (!n) return hipSuccess;
    const uint64_t blocks = (n + 255) / 256;
    const dim3     g((unsigned)(blocks < 4096 ? blocks : 4096)), b(256);
    const auto*    d32 = static_cast<const uint32_t*>(dict);
    const auto*    d64 = static_cast<const uint64_t*>(dict);
    switch (W) {
    case 8: hipLaunchKernelGGL((k_drle_values<uint8_t, uint32_t>), g, b, 0, s, codes, n, d32, dict_n, rv, err); break;
    case 16:
        hipLaunchKernelGGL((k_drle_values<uint16_t, uint32_t>), g, b, 0, s, codes, n, d32, dict_n,
                           reinterpret_cast<uint16_t*>(rv), err);
        break;
    case 32:
        hipLaunchKernelGGL((k_drle_values<uint32_t, uint32_t>), g, b, 0, s, codes, n, d32, dict_n,
                           reinterpret_cast<uint32_t*>(rv), err);
        break;
    case 64:
        hipLaunchKernelGGL((k_drle_values<uint64_t, uint64_t>), g, b, 0, s, codes, n, d64, dict_n,
                           reinterpret_cast<uint64_t*>(rv), err);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_page_dict(unsigned W, int var, const uint8_t* page, const PageMeta& pm, uint32_t vec_c,
                            const void* dict, uint32_t dict_n, uint8_t* out, uint32_t* err, hipStream_t s,
                            const SelArgs* sel, const HzlSrc* hz) {
    if (!vec_c) return hipSuccess;
    switch (W) {
    case 8: return page_dict_w<8>(var, page, pm, vec_c, dict, dict_n, out, err, sel, hz, s);
    case 16: return page_dict_w<16>(var, page, pm, vec_c, dict, dict_n, out, err, sel, hz, s);
    case 32: return page_dict_w<32>(var, page, pm, vec_c, dict, dict_n, out, err, sel, hz, s);
    case 64: return page_dict_w<64>(var, page, pm, vec_c, dict, dict_n, out, err, sel, hz, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_drle_fused(unsigned W, const uint8_t* enc, const HzlChunk* runs, uint32_t n_runs,
                             const HzlChunk* dict, uint32_t n_dict, uint32_t dict_n, uint8_t* rv, uint32_t* err,
                             hipStream_t s) {
    if (!n_runs) return hipSuccess;
    const dim3 g(n_runs < 65535 ? n_runs : 65535), b(256);
#define L(T, D) \
    hipLaunchKernelGGL((k_drle_fused<T, D>), g, b, 0, s, enc, runs, n_runs, dict, n_dict, dict_n, reinterpret_cast<T*>(rv), err)
    switch (W) {
    case 8: L(uint8_t, uint32_t); break;
    case 16: L(uint16_t, uint32_t); break;
    case 32: L(uint32_t, uint32_t); break;
    case 64: L(uint64_t, uint64_t); break;
    default: return hipErrorInvalidValue;
    }
#undef L
    return hipGetLastError();
}

hipError_t launch_hzl(unsigned DW, const uint8_t* enc, const HzlChunk* chunks, uint32_t n_chunks, uint8_t* dict,
                      bool exceptions, hipStream_t s) {
    if (!n_chunks) return hipSuccess;
    const dim3 g(n_chunks < 65535 ? n_chunks : 65535), b(256), ge(n_chunks < 1024 ? n_chunks : 1024);
    if (DW == 32) {
        hipLaunchKernelGGL((k_hzl_chunks<32>), g, b, 0, s, enc, chunks, n_chunks, dict);
        if (exceptions) hipLaunchKernelGGL((k_hzl_exceptions<32>), ge, b, 0, s, enc, chunks, n_chunks, dict);
    } else {
        hipLaunchKernelGGL((k_hzl_chunks<64>), g, b, 0, s, enc, chunks, n_chunks, dict);
        if (exceptions) hipLaunchKernelGGL((k_hzl_exceptions<64>), ge, b, 0, s, enc, chunks, n_chunks, dict);
    }
    return hipGetLastError();
}

}  // namespace fls
