// rowgroup.hpp -- work-list tiles of the persistent row-group decode (k_rowgroup.hip) and their host-side checks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace fls {

// Tile kinds decoded by the persistent kernel. Pages of other expressions (patched FFOR, plain RLE positions,
// DICT_RLE, S_PATCH_REC) and pages whose per-vector metadata is not at fixed strides decode through their
// per-page kernels before the persistent launch.
enum RgKind : uint16_t {
    RG_UNC8, RG_UNC16, RG_UNC32, RG_UNC64,          // UNCOMPRESSED
    RG_FFOR8, RG_FFOR16, RG_FFOR32, RG_FFOR64,      // FFOR / NO_PATCH
    RG_DELTA8, RG_DELTA16, RG_DELTA32, RG_DELTA64,  // DELTA / NO_PATCH
    RG_RLE8, RG_RLE16, RG_RLE32, RG_RLE64,          // RLE / DELTA_NO_PATCH positions
    RG_DICT1_8, RG_DICT1_16, RG_DICT1_32, RG_DICT1_64,  // dictionary, UNCOMPRESSED codes
    RG_DICT2_8, RG_DICT2_16, RG_DICT2_32, RG_DICT2_64,  // dictionary, FFOR codes
    RG_DICT7_8, RG_DICT7_16, RG_DICT7_32, RG_DICT7_64,  // dictionary, DELTA codes
    RG_KIND_COUNT
};

// Vectors [v0, v0 + nv) of one page (nv <= 64: a wave loads the tile's per-vector metadata one vector per lane).
// 64 bytes, read by the kernel with scalar loads.
struct RgTile {
    uint64_t page;    // byte offset of the page in the row-group's encoded buffer
    uint64_t out;     // byte offset of the page's first value in the output buffer
    uint64_t dict;    // device address of the page's decoded dictionary (dictionary kinds), else 0
    uint32_t v0;      // first vector of the tile, page-relative
    uint32_t dict_n;  // dictionary entries
    uint16_t nv;      // vectors in the tile, 1..64
    uint16_t kind;    // RgKind
    uint16_t arr_c;   // buffers of the page
    uint16_t pad;
    uint32_t off[5];  // buffer offsets from the page start
    uint32_t vec_c;   // vectors in the page
};
static_assert(sizeof(RgTile) == 64, "RgTile is read as 64 bytes");

constexpr unsigned kRgMaxTileVec = 64;    // vectors per tile (one metadata lane each)
constexpr unsigned kRgRunsLds    = 2048;  // RLE run values staged in LDS when they fit (bytes per vector)
constexpr unsigned kRgWaves      = 4;     // waves per block

// LDS bytes one wave needs for a kind: the packed rows of one vector plus one pad row (the funnel shift of the
// last value reads the row after the packed data; its bits are masked off), and for RLE the run values.
unsigned rg_wave_lds(int kind);
// Element width (bits) of a kind's decoded values.
unsigned rg_kind_bits(int kind);

// Host check of a tile list against the pages it refers to (run before every upload): every tile names a known
// kind, 1..64 vectors inside its page, buffer offsets inside the page, an output range inside the row-group's
// output, and a dictionary when its kind needs one. Returns "" or the first violation.
struct RgPageExtent {
    uint64_t page, page_size;  // byte range of the page in the encoded buffer
    uint64_t out, out_bytes;   // byte range of the page's decoded values
};
std::string rg_validate_tiles(const std::vector<RgTile>& tiles, const std::vector<RgPageExtent>& pages,
                              uint64_t enc_size, uint64_t out_size);

hipError_t launch_rowgroup(const uint8_t* enc, uint8_t* out, const RgTile* tiles, uint32_t n_tiles, uint32_t* counter,
                           uint32_t* err, unsigned wave_lds, unsigned blocks, hipStream_t s);

}  // namespace fls
