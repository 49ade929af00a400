// rowgroup.hpp -- work-list tiles of the persistent row-group decode.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fls {

// Tile kinds decoded by the persistent kernel (others fall back to their per-page kernels).
enum RgKind : uint16_t {
    RG_FFOR8, RG_FFOR16, RG_FFOR32, RG_FFOR64,
    RG_DELTA8, RG_DELTA16, RG_DELTA32, RG_DELTA64,
    RG_UNC8, RG_UNC16, RG_UNC32, RG_UNC64,
    RG_RLE8, RG_RLE16, RG_RLE32, RG_RLE64,
    RG_DICT8, RG_DICT16, RG_DICT32, RG_DICT64,
    RG_KIND_COUNT
};

// A tile: vectors [v0, v0 + nv) of one page.
struct RgTile {
    uint64_t page;    // byte offset of the page in the row-group's encoded buffer
    uint64_t out;     // byte offset of the page's first value in the output buffer
    uint64_t dict;    // device address of the page's decoded dictionary (dictionary kinds)
    uint32_t dict_n;  // dictionary entries
    uint32_t v0, nv;  // vector range within the page
    uint32_t vec_c;   // vectors in the page
    uint16_t kind;
    uint16_t arr_c;
    uint32_t pad;
};

unsigned   rg_stage_bytes(int kind);
hipError_t launch_rowgroup(const uint8_t* enc, uint8_t* out, const RgTile* tiles, uint32_t n_tiles, uint32_t* counter,
                           uint32_t* err, unsigned lds, unsigned blocks, hipStream_t s);

}  // namespace fls
