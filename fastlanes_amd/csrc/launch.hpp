// launch.hpp -- host-side launchers for the decode kernels and the launch-shape tuning table.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fls {

struct HzlChunk;
struct PageMeta;
struct SelArgs;
struct HzlSrc;
struct RunSrc;

// Kernel families with separately tuned launch shapes.
enum Family : int {
    FAM_STREAM = 0,   // raw FFOR / unpack stream (tuned per width and bit width)
    FAM_PAGE_UNC,     // UNCOMPRESSED pages
    FAM_PAGE_FFOR,    // FFOR pages
    FAM_PAGE_DELTA,   // DELTA pages
    FAM_PAGE_RLE,     // RLE pages (both variants)
    FAM_PAGE_DICT,    // dictionary pages (all variants)
    FAM_ROWGROUP,     // row-group launch: block = KiB of decoded values per tile, max_blocks = blocks per compute unit
    FAM_COUNT
};

// Launch shape: threads per block (multiple of 64, <= 512), non-temporal stores, and a cap on the grid
// (0 = one block per block-sized group of vectors; otherwise blocks stride over the vectors).
struct Tune {
    uint16_t block      = 256;
    uint16_t nt         = 1;
    uint32_t max_blocks = 0;
};
Tune tune_for(int family, unsigned W, unsigned bw);
void tune_set(int family, unsigned W, unsigned bw, Tune t);  // bw 255 = every bit width
int  tune_load(const char* path);                             // returns entries read, <0 on error

inline unsigned grid_for(uint64_t n_vec, const Tune& t, unsigned tpv) {
    const uint64_t vpb    = t.block / tpv;
    uint64_t       blocks = (n_vec + vpb - 1) / vpb;
    if (t.max_blocks && blocks > t.max_blocks) blocks = t.max_blocks;
    if (blocks > 0x7fffffffull) blocks = 0x7fffffffull;
    return (unsigned)(blocks ? blocks : 1);
}

hipError_t launch_unffor_stream(unsigned W, unsigned bw, const uint8_t* packed, const void* bases, uint8_t* out,
                                uint64_t n_vec, hipStream_t s);
hipError_t launch_page_int(unsigned W, int mode, const uint8_t* page, const PageMeta& pm, uint32_t vec_c, uint8_t* out,
                           hipStream_t s, const SelArgs* sel = nullptr);
// runs: where the run values are (DICT_RLE pages, chained run values); nullptr: the page's buffer 0.
// pos_bw1: every vector's DELTA positions have bit width 0 or 1 (8/16-bit values then take the one-lane-per-thread
// kernel)
hipError_t launch_page_rle(unsigned W, int var, const uint8_t* page, const PageMeta& pm, uint32_t vec_c, uint8_t* out,
                           uint32_t* err, hipStream_t s, const SelArgs* sel = nullptr, const RunSrc* runs = nullptr,
                           bool pos_bw1 = false);
// DICT_RLE: rv[i] = dict[codes[i]] (W-bit values from a max(32, W)-bit dictionary)
hipError_t launch_drle_values(unsigned W, const uint32_t* codes, uint64_t n, const void* dict, uint32_t dict_n,
                              uint8_t* rv, uint32_t* err, hipStream_t s);
// DICT_RLE run values straight from the run-code and dictionary HZL chains (neither has exceptions)
hipError_t launch_drle_fused(unsigned W, const uint8_t* enc, const HzlChunk* runs, uint32_t n_runs,
                             const HzlChunk* dict, uint32_t n_dict, uint32_t dict_n, uint8_t* rv, uint32_t* err,
                             hipStream_t s);
// hz: the page's HZL dictionary payload; the launcher decodes it into `dict` (or, when it fits LDS, lets the page
// kernel decode it per block). nullptr: `dict` already holds the decoded dictionary.
hipError_t launch_page_dict(unsigned W, int var, const uint8_t* page, const PageMeta& pm, uint32_t vec_c,
                            const void* dict, uint32_t dict_n, uint8_t* out, uint32_t* err, hipStream_t s,
                            const SelArgs* sel = nullptr, const HzlSrc* hz = nullptr);
// Patched FFOR-shaped pages. var: patch kind in FFOR numbering (2 LL_PATCH, 3 S_PATCH, 4 B_PATCH); rsum: lane bases
// and RSUM (DELTA forms); dict: the decoded dictionary of a dictionary page (u32 codes, values dict[code]), nullptr
// for FFOR / DELTA pages of 32/64-bit values; exc: the exception values (decoded chain of a *_REC form), nullptr:
// the page's exception buffer.
hipError_t launch_page_patch(unsigned W, int var, bool rsum, const void* dict, uint32_t dict_n, const uint8_t* page,
                             const PageMeta& pm, uint32_t vec_c, uint8_t* out, uint32_t* err, hipStream_t s,
                             const SelArgs* sel = nullptr, const uint8_t* exc = nullptr);
hipError_t launch_hzl(unsigned DW, const uint8_t* enc, const HzlChunk* chunks, uint32_t n_chunks, uint8_t* dict,
                      bool exceptions, hipStream_t s);  // exceptions: some chunk has them (else that pass is skipped)

// per-width instantiation units (k_page_w*.hip)
template <int W>
hipError_t page_int_w(int mode, const uint8_t* page, const PageMeta& pm, uint32_t vec_c, uint8_t* out,
                      const SelArgs* sel, hipStream_t s);
template <int W>
hipError_t page_patch_w(int var, bool rsum, const void* dict, uint32_t dict_n, const uint8_t* page, const PageMeta& pm,
                        uint32_t vec_c, uint8_t* out, uint32_t* err, const SelArgs* sel, const uint8_t* exc,
                        hipStream_t s);
template <int W>
hipError_t page_patch_dict_w(int var, bool rsum, const void* dict, uint32_t dict_n, const uint8_t* page,
                             const PageMeta& pm, uint32_t vec_c, uint8_t* out, uint32_t* err, const SelArgs* sel,
                             const uint8_t* exc, hipStream_t s);
template <int W>
hipError_t page_rle_w(int var, const uint8_t* page, const PageMeta& pm, uint32_t vec_c, uint8_t* out, uint32_t* err,
                      const SelArgs* sel, const RunSrc* runs, bool pos_bw1, hipStream_t s);
template <int W>
hipError_t page_dict_w(int var, const uint8_t* page, const PageMeta& pm, uint32_t vec_c, const void* dict,
                       uint32_t dict_n, uint8_t* out, uint32_t* err, const SelArgs* sel, const HzlSrc* hz,
                       hipStream_t s);

}  // namespace fls
