// format.hpp -- host-side view of FastLanes encoded columns: page parsing, validation and the shared constants.
//
// A column is a sequence of page entries; an entry is an optional dictionary page followed by a data page.
//   data page  (src/cor/eng/compressor.cpp:68-86, src/cor/lyt/page/page.cpp:63-130)
//     PageHdr {exp_t u8, var u8, 0 u16, vec_c u16, next u16}   (include/fls/cor/lyt/page/page_hdr.hpp:8-20)
//     arr_c x u32 buffer offsets | (vec_c-1) x arr_c x u32 entry points | buffers
//     `next` (0 in the reference) is 1 when another page entry follows: the reference's u16 vec_c caps a page
//     at 65535 vectors, so larger columns span several pages.
//   dictionary page: dict_hdr_t {exp_t u8, 0, 0, version u8, entry_c u32} (include/fls/cor/lyt/dic/dic.hpp:32-42)
//     DICT: HZL payload of the values (src/cor/eng/hzl_cmp.cpp:43-176); FSST: fsst_export symbol table
//     (src/cor/prm/fsst/libfsst.cpp:582-615).
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace fls {

// ExpT (include/fls/cor/exp/exp_type.hpp:7-28)
enum : uint8_t {
    EXP_UNCOMPRESSED = 2,
    EXP_FFOR         = 4,
    EXP_DELTA        = 5,
    EXP_RLE          = 6,
    EXP_FSST         = 7,
    EXP_DICT         = 8,
    EXP_DICT_FRQ     = 9,
    EXP_DICT_VAL     = 10,
    EXP_DICT_RLE     = 11,
};
// data_t (include/fastlanes.h)
enum : uint8_t {
    DT_DOUBLE = 1,
    DT_I8     = 2,
    DT_I16    = 3,
    DT_I32    = 4,
    DT_I64    = 5,
    DT_U8     = 6,
    DT_U16    = 7,
    DT_U32    = 8,
    DT_U64    = 9,
    DT_STR    = 10,
};

constexpr uint64_t kDataPageMax   = 8ull * 64 * 256 * 1024;  // CFG::PQ::DATA_PG_SZ, include/fls/cfg/cfg.hpp:82
constexpr unsigned kMaxVecPerPage = 65535;                    // u16 PageHdr::vec_c
constexpr unsigned kExcLimit      = 100;                      // CFG::CMPR::EXC_LIMIT_C, include/fls/cfg/cfg.hpp:67
// FFOR variants (include/fls/cor/exp/ffor/ffor_exp.hpp var_t): 1 NO_PATCH, 2 LL_PATCH, 3 S_PATCH, 4 B_PATCH
constexpr uint8_t kFforLL = 2, kFforS = 3, kFforB = 4;
// S_PATCH_REC (var 5): S_PATCH whose exception values are one HZL chain of u64 values per page
// (include/fls/cor/exp/ffor/ffor_exp.hpp:87-94, src/cor/eng/hzl_cmp_fac.cpp:24-31); 64-bit integer columns
constexpr uint8_t kFforSRec = 5;

unsigned dtype_bits(uint8_t dt);
inline bool is_dict_exp(uint8_t e) {
    return e == EXP_DICT || e == EXP_DICT_FRQ || e == EXP_DICT_VAL || e == EXP_DICT_RLE;
}
// DICT_RLE var DELTA_NO_PATCH_REC (include/fls/cor/exp/dict_rle/dict_rle_exp.hpp:24-30): the only DICT_RLE variant
// in the integer pools (src/cor/exp_pool.cpp:53, id 27)
constexpr uint8_t kDictRleRec = 4;

struct Expr {
    uint8_t exp_t, var, arr_c;
    bool    dict;
};
// Pool ids (src/cor/exp_pool.cpp:23-100) -> expression; false if not supported.
bool expr_from_id(uint8_t dtype, uint8_t id, Expr& e);
bool expr_from_page(uint8_t exp_t, uint8_t var, Expr& e);
// Whether a column of this type may hold pages of the expression: patches resolve for 32/64-bit values only
// (src/cor/prm/patch/*/..._compress.cpp ResolveCompressFunc), chains of THzlCmp<uint64_t> hold 64-bit values here,
// doubles take UNCOMPRESSED and the RLE forms (their pool's dictionaries are ALP-coded).
bool expr_fits_type(const Expr& e, uint8_t dtype);

// Variant numbering of DELTA (include/fls/cor/exp/delta/delta_exp.hpp:24-30), RLE
// (include/fls/cor/exp/rle/rle_exp.hpp:24-30) and DICT / DICT_FRQ / DICT_VAL
// (include/fls/cor/exp/dict/dict_exp.hpp:30-42).
constexpr uint8_t kDeltaLL = 2, kDeltaS = 3, kDeltaB = 4, kDeltaBRec = 5;
constexpr uint8_t kRleRec = 3, kRleDeltaRec = 4, kRleDeltaRec2 = 5;
constexpr uint8_t kDictFforLL = 3, kDictFforS = 4, kDictFforB = 5, kDictFforSRec = 6, kDictDelta = 7, kDictDeltaLL = 8,
                  kDictDeltaS = 9, kDictDeltaB = 10, kDictDeltaBRec = 11;

// The FFOR-shaped pages -- FFOR, DELTA and the dictionary code pages other than UNCOMPRESSED -- lay a vector out as
// [lane bases (RSUM)] [exceptions][positions | bitmap | count+first (patch)] [bw][base][packed]; these describe it.
// Patch kind in FFOR numbering (kFforLL / kFforS / kFforB), 0 without a patch step.
uint8_t expr_patch(const Expr& e);
// Values pass RSUM after FFOR (and after the patch): DELTA, dictionary DELTA_* variants.
bool expr_rsum(const Expr& e);
// Buffers that a page stores as one HZL chain each (RPN is_compressed flags): bit b = buffer b.
unsigned expr_rec(const Expr& e);
// Width of those chains' values: THzlCmp<idx_t> for dictionary expressions, THzlCmp<uint64_t> otherwise
// (src/cor/eng/hzl_cmp_fac.cpp:15-35).
unsigned expr_rec_width(const Expr& e);
// FFOR-shaped page (see above).
bool expr_ffor_shaped(const Expr& e);

struct HzlChunkH {
    uint64_t src, dst, base;
    uint32_t bw, exc_c, n;
};

struct FsstTableH {
    uint8_t  len[255];
    uint64_t symbol[255];
};

struct PageEntry {
    uint64_t dict_off  = ~0ull;  // dictionary page offset, if any
    uint64_t page_off  = 0;      // data page offset
    uint64_t page_size = 0;      // bytes of the data page
    uint64_t out_index = 0;      // first decoded value
    uint32_t vec_c = 0, arr_c = 0;
    uint8_t  exp_t = 0, var = 0;
    uint32_t boff[8]    = {};     // buffer offsets from the page start
    bool     fixed_meta = false;  // every fixed-size buffer (bw, base, lane bases, plain values) sits at v * size
    // dictionary (integer columns)
    uint32_t               dict_n = 0;
    std::vector<HzlChunkH> hzl;
    // Buffers stored as one HZL chain per page (expr_rec): buffer 0 (DICT_RLE run codes, S_PATCH_REC exception
    // values, RLE run values) and buffer 1 (exceptions of the DELTA B_PATCH_REC forms, RLE position lane bases);
    // their chunks and decoded value counts (in values of expr_rec_width bits)
    std::vector<HzlChunkH> run_hzl;
    uint64_t               run_codes = 0;
    std::vector<HzlChunkH> rec1_hzl;
    uint64_t               rec1_n = 0;
    uint8_t                rec = 0, rec_width = 0;
    // RLE / DICT_RLE with DELTA positions: the largest position bit width over the page's vectors (0 or 1 for
    // every page the encoders write)
    uint8_t pos_bw_max = 0;
    // FSST (string columns)
    FsstTableH fsst;
};

struct Column {
    uint8_t                dtype = 0;
    uint64_t               n_values = 0, size = 0;
    std::vector<PageEntry> pages;
};

// Parses and validates an encoded column (every offset, entry point, bit width and extent is checked so that
// device kernels only touch bytes inside [0, size)). Returns an empty string on success, else the reason.
std::string parse_column(const uint8_t* p, uint64_t size, uint8_t dtype, Column& out);
// Size of the page-entry chain starting at p when the total size is not known (bounded by `limit`).
std::string measure_column(const uint8_t* p, uint64_t limit, uint8_t dtype, uint64_t& size);

// Vector v's slice of buffer b (host view).
inline uint64_t ep_of(const uint8_t* page, unsigned arr_c, unsigned v, unsigned b) {
    uint32_t ep = 0;
    if (v) std::memcpy(&ep, page + 8 + 4ull * arr_c + 4ull * ((uint64_t)(v - 1) * arr_c + b), 4);
    return ep;
}
inline uint64_t off_of(const uint8_t* page, unsigned b) {
    uint32_t o;
    std::memcpy(&o, page + 8 + 4ull * b, 4);
    return o;
}

// Encoder (product side). Returns the encoded size, or -1 unsupported / -2 capacity.
int64_t encode_column(uint8_t dtype, uint8_t expr_id, const void* in, uint64_t n, uint8_t* out, uint64_t cap,
                      uint64_t* metadata_size);
int64_t encode_str_column(uint8_t expr_id, const uint32_t* offsets, const uint8_t* bytes, uint64_t n, uint8_t* out,
                          uint64_t cap);

}  // namespace fls
