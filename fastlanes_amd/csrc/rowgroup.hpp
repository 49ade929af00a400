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
// The tile's input bytes are up to three contiguous ranges of its page, staged whole into the wave's LDS slot:
//   integer pages   0: packed rows (UNCOMPRESSED: the values)      1: DELTA lane bases
//   RLE pages       0: packed u16 position rows                    1: run values        2: position lane bases
//   dictionary      0: packed u32 code rows (UNCOMPRESSED: codes)  1: DELTA lane bases
// Range 0 is followed in LDS by one pad row of 128 bytes (the funnel shift of a vector's last value reads the row
// after its packed rows; those bits are masked off). 128 bytes, read by the kernel with scalar loads.
struct RgTile {
    uint64_t page;     // byte offset of the page in the row-group's encoded buffer
    uint64_t out;      // byte offset of the page's first value in the output buffer
    uint64_t dict;     // device address of the page's decoded dictionary (dictionary kinds), else 0
    uint32_t v0;       // first vector of the tile, page-relative
    uint32_t dict_n;   // dictionary entries
    uint32_t vec_c;    // vectors in the page
    uint16_t nv;       // vectors in the tile, 1..64
    uint16_t kind;     // RgKind
    uint16_t arr_c;    // buffers of the page
    uint16_t n_rng;    // staged ranges, 1..3
    uint32_t off[5];   // buffer offsets from the page start
    uint32_t rsrc[3];  // staged ranges: byte offset from the page start
    uint32_t rlen[3];  //                bytes
    uint32_t rlds[3];  //                byte offset in the wave's LDS slot (16-aligned)
    uint32_t chunks;   // 16-byte chunks over all ranges
    // chunk c (0 <= c < chunks) of range r is page byte 16 c + sdelta[r] and slot byte 16 c + ldelta[r]; range 1
    // starts at chunk ceil(rlen[0] / 16), range 2 at ceil(rlen[0] / 16) + ceil(rlen[1] / 16)
    int32_t  sdelta[3];
    int32_t  ldelta[3];
};
static_assert(sizeof(RgTile) == 128, "RgTile is read as 128 bytes");


constexpr unsigned kRgMaxTileVec = 64;    // vectors per tile (one metadata lane each)
constexpr unsigned kRgWaves      = 4;     // waves per block
constexpr unsigned kRgPrefetch   = 8;     // 16-byte chunks per lane held in registers ahead of a tile's decode
constexpr unsigned kRgStage      = 1024 * kRgPrefetch;  // staged bytes the registers carry (larger tiles copy the
                                                        // rest when they are staged)
constexpr unsigned kRgDump       = 16;    // LDS bytes at the end of a wave slot for the lanes without a chunk

// LDS bytes of a wave's output buffer for W-bit values: one decoded vector (128 W bytes) with 16 pad bytes after
// every 256 (the threads' 16-byte words land on different banks), read back lane-consecutively for the stores.
constexpr unsigned rg_out_lds(unsigned W) { return 128 * W + 8 * W; }

// Element width (bits) of a kind's decoded values.
unsigned rg_kind_bits(int kind);

// Host check of a tile list against the pages it refers to (run before every upload): every tile names a known
// kind, 1..64 vectors inside its page, buffer offsets and staged ranges inside the page, ranges that fit the
// wave's LDS slot, an output range inside the row-group's output, and a dictionary when its kind needs one.
// Returns "" or the first violation.
struct RgPageExtent {
    uint64_t page, page_size;  // byte range of the page in the encoded buffer
    uint64_t out, out_bytes;   // byte range of the page's decoded values
};
std::string rg_validate_tiles(const std::vector<RgTile>& tiles, const std::vector<RgPageExtent>& pages,
                              uint64_t enc_size, uint64_t out_size, unsigned wave_lds);

// Orders the tiles into one contiguous run per wave of a launch of `waves` waves and returns the runs' bounds
// (waves + 1 entries): longest tiles first, each to the wave with the least estimated work so far (decoded plus
// staged bytes, plus a fixed cost per vector); within a run, tiles keep their list order.
std::vector<uint32_t> rg_assign_waves(std::vector<RgTile>& tiles, uint32_t waves);

// Blocks of `waves` waves one compute unit holds at once with `slot_lds` LDS bytes per wave.
unsigned rg_blocks_per_cu(unsigned slot_lds, unsigned waves);

// One launch: `blocks` blocks of `waves` waves (1..kRgWaves), wave w decoding tiles [wave_begin[w],
// wave_begin[w+1]). Each wave's LDS slot holds its staged ranges (`wave_lds` bytes) and then its output buffer
// (`out_lds` bytes).
hipError_t launch_rowgroup(const uint8_t* enc, uint8_t* out, const RgTile* tiles, const uint32_t* wave_begin,
                           uint32_t* err, unsigned wave_lds, unsigned out_lds, unsigned blocks, unsigned waves,
                           hipStream_t s);

}  // namespace fls
