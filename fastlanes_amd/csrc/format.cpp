// format.cpp -- page parsing/validation and the host encoder for the FastLanes column format.
#include "format.hpp"

#include <algorithm>
#include <thread>
#include <unordered_map>

namespace fls {

unsigned dtype_bits(uint8_t dt) {
    switch (dt) {
    case DT_I8:
    case DT_U8: return 8;
    case DT_I16:
    case DT_U16: return 16;
    case DT_I32:
    case DT_U32: return 32;
    case DT_I64:
    case DT_U64:
    case DT_DOUBLE: return 64;
    default: return 0;
    }
}

bool expr_from_page(uint8_t exp_t, uint8_t var, Expr& e) {
    auto set = [&](uint8_t arr_c, bool dict) { return e = {exp_t, var, arr_c, dict}, true; };
    switch (exp_t) {
    case EXP_UNCOMPRESSED: return var == 1 && set(1, false);
    case EXP_FFOR:
        if (var == 1) return set(3, false);
        return (var == kFforLL || var == kFforS || var == kFforB || var == kFforSRec) && set(5, false);
    case EXP_DELTA:
        if (var == 1) return set(4, false);
        return var >= kDeltaLL && var <= kDeltaBRec && set(6, false);
    case EXP_RLE:
        if (var == 1 || var == kRleRec) return set(2, false);
        return (var == 2 || var == kRleDeltaRec || var == kRleDeltaRec2) && set(5, false);
    case EXP_DICT:
    case EXP_DICT_FRQ:
    case EXP_DICT_VAL:
        if (var == 1) return set(1, true);
        if (var == 2) return set(3, true);
        if (var >= kDictFforLL && var <= kDictFforSRec) return set(5, true);
        if (var == kDictDelta) return set(4, true);
        return var >= kDictDeltaLL && var <= kDictDeltaBRec && set(6, true);
    case EXP_DICT_RLE: return var == kDictRleRec && set(5, true);
    case EXP_FSST: return var == 1 && set(2, true);
    default: return false;
    }
}

uint8_t expr_patch(const Expr& e) {
    switch (e.exp_t) {
    case EXP_FFOR: return e.var == kFforSRec ? kFforS : e.var == 1 ? 0 : e.var;
    case EXP_DELTA: return e.var == kDeltaLL ? kFforLL : e.var == kDeltaS ? kFforS : e.var >= kDeltaB ? kFforB : 0;
    case EXP_DICT:
    case EXP_DICT_FRQ:
    case EXP_DICT_VAL:
        switch (e.var) {
        case kDictFforLL:
        case kDictDeltaLL: return kFforLL;
        case kDictFforS:
        case kDictFforSRec:
        case kDictDeltaS: return kFforS;
        case kDictFforB:
        case kDictDeltaB:
        case kDictDeltaBRec: return kFforB;
        default: return 0;
        }
    default: return 0;
    }
}
bool expr_rsum(const Expr& e) {
    return e.exp_t == EXP_DELTA || ((e.exp_t == EXP_DICT || e.exp_t == EXP_DICT_FRQ || e.exp_t == EXP_DICT_VAL) &&
                                    e.var >= kDictDelta);
}
unsigned expr_rec(const Expr& e) {
    switch (e.exp_t) {
    case EXP_FFOR: return e.var == kFforSRec ? 1u : 0u;
    case EXP_DELTA: return e.var == kDeltaBRec ? 2u : 0u;
    case EXP_RLE: return e.var == kRleRec || e.var == kRleDeltaRec ? 1u : e.var == kRleDeltaRec2 ? 3u : 0u;
    case EXP_DICT:
    case EXP_DICT_FRQ:
    case EXP_DICT_VAL: return e.var == kDictFforSRec ? 1u : e.var == kDictDeltaBRec ? 2u : 0u;
    case EXP_DICT_RLE: return 1u;
    default: return 0u;
    }
}
unsigned expr_rec_width(const Expr& e) { return is_dict_exp(e.exp_t) ? 32u : 64u; }
bool     expr_ffor_shaped(const Expr& e) {
    return e.exp_t == EXP_FFOR || e.exp_t == EXP_DELTA ||
           ((e.exp_t == EXP_DICT || e.exp_t == EXP_DICT_FRQ || e.exp_t == EXP_DICT_VAL) && e.var != 1);
}
bool expr_fits_type(const Expr& e, uint8_t dtype) {
    const unsigned W = dtype_bits(dtype);
    if (dtype == DT_STR) return e.exp_t == EXP_FSST;
    if (!W || e.exp_t == EXP_FSST) return false;
    // doubles: UNCOMPRESSED and the RLE forms of the double pool (src/cor/exp_pool.cpp:105-112); its dictionary
    // forms store the dictionary through ALP (src/cor/lyt/dic/dic.cpp: TDic<dbl_pt>::Compress), out of scope here
    if (dtype == DT_DOUBLE) return e.exp_t == EXP_UNCOMPRESSED || e.exp_t == EXP_RLE;
    if (!e.dict && expr_patch(e) && W < 32) return false;            // patched values
    if (!e.dict && expr_rec(e) && W != 64) return false;             // THzlCmp<uint64_t> chains
    return true;
}

bool expr_from_id(uint8_t dtype, uint8_t id, Expr& e) {
    if (dtype == DT_STR) {
        if (id == 13) return e = {EXP_FSST, 1, 2, true}, true;
        return false;
    }
    // (ExpT, var) of the integer pool, src/cor/exp_pool.cpp:26-64
    static const uint8_t pool[39][2] = {
        {EXP_UNCOMPRESSED, 1},        {EXP_FFOR, 1},                 {EXP_FFOR, kFforS},
        {EXP_FFOR, kFforSRec},        {EXP_FFOR, kFforB},            {EXP_FFOR, kFforLL},
        {EXP_RLE, 2},                 {EXP_RLE, 1},                  {EXP_RLE, kRleRec},
        {EXP_RLE, kRleDeltaRec},      {EXP_RLE, kRleDeltaRec2},      {EXP_DELTA, 1},
        {EXP_DELTA, kDeltaS},         {EXP_DELTA, kDeltaB},          {EXP_DELTA, kDeltaLL},
        {EXP_DELTA, kDeltaBRec},      {EXP_DICT, 1},                 {EXP_DICT, 2},
        {EXP_DICT, kDictFforB},       {EXP_DICT, kDictFforS},        {EXP_DICT, kDictFforLL},
        {EXP_DICT, kDictFforSRec},    {EXP_DICT, kDictDelta},        {EXP_DICT, kDictDeltaS},
        {EXP_DICT, kDictDeltaLL},     {EXP_DICT, kDictDeltaB},       {EXP_DICT, kDictDeltaBRec},
        {EXP_DICT_RLE, kDictRleRec},  {EXP_DICT_FRQ, 1},             {EXP_DICT_FRQ, 2},
        {EXP_DICT_FRQ, kDictFforB},   {EXP_DICT_FRQ, kDictFforS},    {EXP_DICT_FRQ, kDictFforLL},
        {EXP_DICT_FRQ, kDictFforSRec}, {EXP_DICT_FRQ, kDictDelta},   {EXP_DICT_FRQ, kDictDeltaS},
        {EXP_DICT_FRQ, kDictDeltaLL}, {EXP_DICT_FRQ, kDictDeltaB},   {EXP_DICT_FRQ, kDictDeltaBRec},
    };
    if (id >= 39 || !expr_from_page(pool[id][0], pool[id][1], e)) return false;
    return expr_fits_type(e, dtype);
}

// ------------------------------------------------------------------------------------------------------------------
// Parsing
// ------------------------------------------------------------------------------------------------------------------
namespace {

struct Reader {
    const uint8_t* p;
    uint64_t       size;
    bool           has(uint64_t off, uint64_t n) const { return off <= size && n <= size - off; }
};

// Parses the HZL chain of a dictionary page's payload starting at `off`; fills chunks; returns bytes consumed.
std::string parse_hzl(const Reader& r, uint64_t off, unsigned dw, uint32_t entry_c, std::vector<HzlChunkH>& out,
                      uint64_t& used) {
    const unsigned sz = dw / 8;
    uint64_t       i = off, n = 0;
    for (;;) {
        if (!r.has(i, 12)) return "dictionary payload truncated";
        const uint8_t bw = r.p[i];
        uint16_t      exc_c;
        std::memcpy(&exc_c, r.p + i + 2, 2);
        uint64_t base = 0;
        std::memcpy(&base, r.p + i + 4, 8);
        i += 12;
        if (bw == 255) {
            const uint64_t bytes = base;
            if (bytes % sz || !r.has(i, bytes)) return "dictionary tail out of bounds";
            out.push_back({i, n, 0, 255, 0, (uint32_t)(bytes / sz)});
            n += bytes / sz;
            i += bytes;
            break;
        }
        if (bw > dw) return "dictionary chunk bit width too large";
        if (!r.has(i, 128ull * bw + (uint64_t)exc_c * (2 + sz))) return "dictionary chunk out of bounds";
        if (dw == 32) base &= 0xffffffffull;
        out.push_back({i, n, base, bw, exc_c, 1024});
        i += 128ull * bw + (uint64_t)exc_c * (2 + sz);
        n += 1024;
        if (n > (1ull << 32)) return "dictionary too large";
    }
    if (entry_c != ~0u && n != entry_c) return "dictionary entry count mismatch";
    used = i - off;
    return "";
}

std::string parse_fsst_table(const Reader& r, uint64_t off, FsstTableH& t, uint64_t& used) {
    // fsst_import (src/cor/prm/fsst/libfsst.cpp:617-652)
    if (!r.has(off, 17)) return "fsst table truncated";
    uint64_t version;
    std::memcpy(&version, r.p + off, 8);
    if ((version >> 32) != 20190218ull) return "fsst table version";
    const bool zt = r.p[off + 8] & 1;
    uint8_t    lh[8];
    std::memcpy(lh, r.p + off + 9, 8);
    unsigned code = zt, pos = 17;
    t.len[0]      = 1;
    t.symbol[0]   = 0;
    if (zt) lh[0]--;
    for (unsigned l = 1; l <= 8; ++l)
        for (unsigned i = 0; i < lh[l & 7]; ++i, ++code) {
            if (code >= 255) return "fsst table too many symbols";
            const unsigned ln = (l & 7) + 1;
            if (!r.has(off + pos, ln)) return "fsst table truncated";
            t.len[code]    = (uint8_t)ln;
            t.symbol[code] = 0;
            std::memcpy(&t.symbol[code], r.p + off + pos, ln);
            pos += ln;
        }
    while (code < 255) t.symbol[code] = 32774747032022883ull, t.len[code++] = 8;  // "corrupt"
    used = pos;
    return "";
}

// Checks a data page and returns its size.
std::string parse_page(const Reader& r, uint64_t off, uint8_t dtype, PageEntry& pe) {
    if (!r.has(off, 8)) return "page header truncated";
    const uint8_t* pg = r.p + off;
    Expr           e;
    if (!expr_from_page(pg[0], pg[1], e)) return "unsupported expression in page header";
    pe.exp_t = pg[0], pe.var = pg[1], pe.arr_c = e.arr_c;
    pe.vec_c = (unsigned)pg[4] | ((unsigned)pg[5] << 8);
    const unsigned W = dtype == DT_STR ? 8 : dtype_bits(dtype);
    if (!W) return "unsupported data type";
    if (dtype == DT_STR && e.exp_t != EXP_FSST) return "string column with integer expression";
    if (dtype != DT_STR && e.exp_t == EXP_FSST) return "integer column with string expression";
    if (!expr_fits_type(e, dtype)) return "expression not defined for this data type";
    const unsigned arr_c = e.arr_c;
    const uint64_t meta  = 8 + 4ull * arr_c + 4ull * arr_c * (pe.vec_c ? pe.vec_c - 1 : 0);
    if (!r.has(off, meta)) return "page metadata truncated";
    if (pe.vec_c == 0) return pe.page_size = 8 + 4ull * arr_c, "";
    uint64_t       boff[8];
    const uint64_t avail = r.size - off;
    for (unsigned b = 0; b < arr_c; ++b) {
        boff[b] = off_of(pg, b);
        if (boff[b] < meta || boff[b] > avail || (b && boff[b] < boff[b - 1])) return "buffer offset out of range";
    }
    // per-vector extents
    const unsigned sz = W / 8;
    auto           ep = [&](unsigned v, unsigned b) { return ep_of(pg, arr_c, v, b); };
    // buffers stored as one HZL chain per page: their entry points address the decoded bytes
    const unsigned rec = expr_rec(e), rdw = expr_rec_width(e);
    pe.rec = (uint8_t)rec, pe.rec_width = (uint8_t)rdw;
    uint64_t rec_bytes[2] = {0, 0};
    for (unsigned b = 0; b < 2; ++b) {
        if (!(rec >> b & 1)) continue;
        auto&    chunks = b ? pe.rec1_hzl : pe.run_hzl;
        uint64_t used = 0;
        auto     err  = parse_hzl(r, off + boff[b], rdw, ~0u, chunks, used);
        if (!err.empty()) return "recursive buffer " + std::to_string(b) + ": " + err;
        if (used != boff[b + 1] - boff[b]) return "recursive buffer " + std::to_string(b) + " does not fill its slot";
        uint64_t n = 0;
        for (const auto& c : chunks) n += c.n;
        (b ? pe.rec1_n : pe.run_codes) = n;
        rec_bytes[b]                   = rdw / 8 * n;
        if (rec_bytes[b] > 0xffffffffull) return "recursive buffer exceeds 32-bit entry points";
    }
    auto cap = [&](unsigned b) {
        return b < 2 && (rec >> b & 1) ? rec_bytes[b] : (b + 1 < arr_c ? boff[b + 1] : avail) - boff[b];
    };
    auto           need = [&](unsigned v, unsigned b, uint64_t n) -> bool {
        const uint64_t s = ep(v, b);
        const uint64_t lim = v + 1 < pe.vec_c ? ep(v + 1, b) : cap(b);
        return s <= lim && lim <= cap(b) && n <= lim - s;
    };
    auto bwc = [&](unsigned v, unsigned b) { return (unsigned)pg[boff[b] + ep(v, b)]; };
    // FFOR-shaped pages: [lane bases][exceptions][positions][bw][base][packed] over CW-bit values
    const bool     ffs   = expr_ffor_shaped(e);
    const uint8_t  patch = expr_patch(e);
    const unsigned CW    = e.dict ? 32 : W, csz = CW / 8;
    const unsigned o     = expr_rsum(e) ? 1 : 0;  // first buffer after the lane bases
    const unsigned q     = o + (patch ? 2 : 0);   // bw buffer
    uint64_t end = 0;  // bytes of the last buffer used by the last vector
    for (unsigned v = 0; v < pe.vec_c; ++v) {
        bool ok = true;
        if (ffs) {
            if (o) ok = need(v, 0, 128);
            ok = ok && need(v, q, 1) && need(v, q + 1, csz);
            if (ok) {
                const unsigned bw = bwc(v, q);
                ok = bw <= CW && need(v, q + 2, 128ull * bw);
            }
            if (ok && patch) {  // [exception values][positions | bitmap | count + first position]
                const uint8_t* b1  = pg + boff[o + 1] + ep(v, o + 1);
                uint32_t       exc = 0;
                if (patch == kFforB) {
                    ok = need(v, o + 1, 128);
                    for (unsigned k = 0; ok && k < 128; ++k) exc += (uint32_t)__builtin_popcount(b1[k]);
                } else {
                    ok = need(v, o + 1, patch == kFforS ? 2 : 4);
                    if (ok) {
                        exc = (uint32_t)b1[0] | ((uint32_t)b1[1] << 8);
                        if (patch == kFforS) {
                            ok = exc <= 1024 && need(v, o + 1, 2 + 2ull * exc);
                            for (uint32_t k = 0; ok && k < exc; ++k)
                                ok = ((uint32_t)b1[2 + 2 * k] | ((uint32_t)b1[3 + 2 * k] << 8)) < 1024;
                        } else {
                            ok = exc <= 1024 && (exc == 0 || ((uint32_t)b1[2] | ((uint32_t)b1[3] << 8)) < 1024);
                        }
                    }
                }
                ok = ok && need(v, o, (uint64_t)exc * csz);
            }
        } else {
            switch (e.exp_t) {
            case EXP_UNCOMPRESSED: ok = need(v, 0, 1024ull * sz); break;
            case EXP_RLE:
            case EXP_DICT_RLE: {  // [run values | run codes][positions: u16 plain, or RSUM + FFOR]
                const unsigned rsz = e.exp_t == EXP_DICT_RLE ? 4 : sz;
                ok = ep(v, 0) % rsz == 0 && need(v, 0, rsz);  // at least one run
                if (e.var == 1 || e.var == kRleRec) ok = ok && need(v, 1, 2048);
                else {
                    ok = ok && need(v, 1, 128) && need(v, 2, 1) && need(v, 3, 2);
                    if (!ok) break;
                    const unsigned bw = bwc(v, 2);
                    ok = bw <= 16 && need(v, 4, 128ull * bw);
                    if (bw > pe.pos_bw_max) pe.pos_bw_max = (uint8_t)bw;
                }
                break;
            }
            case EXP_FSST: ok = need(v, 1, 4096) && need(v, 0, 0); break;
            default: ok = need(v, 0, 4096); break;  // dictionary codes, UNCOMPRESSED
            }
        }
        if (!ok) return "vector " + std::to_string(v) + " exceeds its page buffers";
    }
    for (unsigned b = 0; b < arr_c; ++b) pe.boff[b] = (uint32_t)boff[b];
    {  // fixed per-vector sizes of the metadata buffers (0 = variable size)
        unsigned fsz[8] = {};
        if (ffs) {
            if (o) fsz[0] = 128;
            fsz[q] = 1, fsz[q + 1] = csz;
            if (patch) fsz[o + 1] = patch == kFforB ? 128 : patch == kFforLL ? 4 : 0;  // S: variable
        } else {
            switch (e.exp_t) {
            case EXP_UNCOMPRESSED: fsz[0] = 1024 * sz; break;
            case EXP_RLE:
            case EXP_DICT_RLE:
                if (e.var == 1 || e.var == kRleRec) fsz[1] = 2048;
                else fsz[1] = 128, fsz[2] = 1, fsz[3] = 2;
                break;
            case EXP_FSST: fsz[1] = 4096; break;
            default: fsz[0] = 4096;
            }
        }
        bool fixed = true;
        for (unsigned v = 0; v < pe.vec_c && fixed; ++v)
            for (unsigned b = 0; b < arr_c; ++b)
                if (fsz[b] && ep(v, b) != (uint64_t)v * fsz[b]) fixed = false;
        pe.fixed_meta = fixed;
    }
    // page size: last buffer's start + last vector's slice end
    {
        const unsigned v = pe.vec_c - 1, b = arr_c - 1;
        uint64_t       tail = 0;
        if (ffs) tail = 128ull * bwc(v, b - 2);
        else switch (e.exp_t) {
            case EXP_UNCOMPRESSED: tail = 1024ull * sz; break;
            case EXP_RLE:
            case EXP_DICT_RLE: tail = e.var == 1 || e.var == kRleRec ? 2048 : 128ull * bwc(v, b - 2); break;
            default: tail = 4096; break;  // FSST offsets, dictionary codes
            }
        end = boff[b] + ep(v, b) + tail;
    }
    // FSST: compressed byte extents must cover the offsets' last value (checked in parse_column via host walk).
    pe.page_size = end;
    return "";
}

}  // namespace

std::string parse_column(const uint8_t* p, uint64_t size, uint8_t dtype, Column& out) {
    Reader r {p, size};
    out       = Column {};
    out.dtype = dtype;
    out.size  = size;
    const unsigned W  = dtype == DT_STR ? 8 : dtype_bits(dtype);
    const unsigned dw = W < 32 ? 32 : W;
    if (!W) return "unsupported data type";
    uint64_t pos = 0;
    for (;;) {
        PageEntry pe;
        if (!r.has(pos, 8)) return "truncated column";
        if (is_dict_exp(p[pos]) || p[pos] == EXP_FSST) {
            pe.dict_off = pos;
            uint64_t used = 0;
            if (p[pos] == EXP_FSST) {
                if (dtype != DT_STR) return "FSST dictionary on a non-string column";
                auto err = parse_fsst_table(r, pos + 8, pe.fsst, used);
                if (!err.empty()) return err;
            } else {
                if (dtype == DT_STR || dtype == DT_DOUBLE) return "dictionary on an unsupported type";
                uint32_t n;
                std::memcpy(&n, p + pos + 4, 4);
                pe.dict_n = n;
                auto err  = parse_hzl(r, pos + 8, dw, n, pe.hzl, used);
                if (!err.empty()) return err;
            }
            pos += 8 + used;
        }
        pe.page_off = pos;
        auto err    = parse_page(r, pos, dtype, pe);
        if (!err.empty()) return err;
        if (is_dict_exp(pe.exp_t) != (pe.dict_off != ~0ull && p[pe.dict_off] != EXP_FSST) ||
            (pe.exp_t == EXP_FSST) != (pe.dict_off != ~0ull && p[pe.dict_off] == EXP_FSST))
            return "dictionary page does not match the data page's expression";
        if (is_dict_exp(pe.exp_t) && pe.vec_c && pe.dict_n == 0) return "empty dictionary";
        if (pe.exp_t == EXP_FSST) {
            // every vector's compressed bytes must lie inside buffer 0's slice: end offsets are monotone and
            // bounded by the slice; checked here on the host copy.
            const uint8_t* pg = p + pos;
            for (unsigned v = 0; v < pe.vec_c; ++v) {
                const uint64_t s0 = ep_of(pg, 2, v, 0);
                const uint64_t l0 = (v + 1 < pe.vec_c ? ep_of(pg, 2, v + 1, 0) : off_of(pg, 1) - off_of(pg, 0)) - s0;
                const uint8_t* ends = pg + off_of(pg, 1) + ep_of(pg, 2, v, 1);
                uint32_t       mx   = 0;
                for (unsigned k = 0; k < 1024; ++k) {
                    uint32_t e;
                    std::memcpy(&e, ends + 4 * k, 4);
                    mx = e > mx ? e : mx;
                }
                if (mx > l0) return "fsst vector offsets exceed its bytes";
            }
        }
        pe.out_index = out.n_values;
        out.n_values += 1024ull * pe.vec_c;
        const bool more = p[pos + 6] != 0;
        out.pages.push_back(std::move(pe));
        pos += out.pages.back().page_size;
        if (!more) break;
        if (out.pages.size() > (1u << 20)) return "too many pages";
    }
    if (pos > size) return "column larger than the given size";
    out.size = pos;
    return "";
}

std::string measure_column(const uint8_t* p, uint64_t limit, uint8_t dtype, uint64_t& size) {
    Column c;
    auto   err = parse_column(p, limit, dtype, c);
    size       = err.empty() ? c.size : 0;
    return err;
}

// ------------------------------------------------------------------------------------------------------------------
// Encoder: per-vector compress RPNs of include/fls/cor/exp/*/*_exp.hpp, pages as TCompressor::FlushPage writes them.
// ------------------------------------------------------------------------------------------------------------------
namespace {

constexpr unsigned kOrder8[8] = {0, 4, 2, 6, 1, 5, 3, 7};
inline unsigned orig_of(unsigned k) { return (k >> 7) + 8 * kOrder8[(k >> 4) & 7] + 64 * (k & 15); }

template <typename T>
void transpose(const T* in, T* out) {
    for (unsigned k = 0; k < 1024; ++k) out[k] = in[orig_of(k)];
}
template <typename T>
constexpr unsigned bits() {
    return 8 * sizeof(T);
}
template <typename T>
unsigned lane_pos(unsigned i, unsigned s) {
    constexpr unsigned W = bits<T>(), L = 1024 / W;
    static const unsigned o1[1] = {0}, o2[2] = {0, 1}, o4[4] = {0, 2, 1, 3};
    const unsigned*       ord = W == 8 ? o1 : W == 16 ? o2 : W == 32 ? o4 : kOrder8;
    return i + 128 * (s % 8) + L * ord[s / 8];
}
template <typename T>
void unrsum(const T* in, T* out) {  // generated unrsum: first step 0, then lane-order differences
    constexpr unsigned W = bits<T>(), L = 1024 / W;
    for (unsigned i = 0; i < L; ++i) {
        out[lane_pos<T>(i, 0)] = 0;
        for (unsigned s = 1; s < W; ++s) out[lane_pos<T>(i, s)] = (T)(in[lane_pos<T>(i, s)] - in[lane_pos<T>(i, s - 1)]);
    }
}
// ANALYZE (src/cor/prm/ffor/ffor_analyze.cpp:78-102): base = min, bw = exact bit length of max - min. `sgn`: min and
// max as signed values, as ffor_prm::ResolveAnalyzeFunc resolves it for int64 columns (ffor_analyze.cpp:145-160);
// every other type compares unsigned.
template <typename T>
uint8_t analyze(const T* in, T& base, bool sgn = false) {
    const T bias = sgn ? (T)((T)1 << (8 * sizeof(T) - 1)) : (T)0;
    T       mn = (T)(in[0] ^ bias), mx = mn;
    for (unsigned i = 1; i < 1024; ++i) {
        const T x = (T)(in[i] ^ bias);
        mn = x < mn ? x : mn, mx = x > mx ? x : mx;
    }
    base             = (T)(mn ^ bias);
    const uint64_t d = (uint64_t)(T)(mx - mn);
    return (uint8_t)(d ? 64 - __builtin_clzll(d) : 0);
}
// Patched FFOR parameters of one transposed vector, as ffor_patch_analyze / find_best_params choose them
// (src/cor/prm/ffor/ffor_analyze.cpp:41-75, :105-140): every distinct value is tried as the base with every bit width
// below W; the values outside [base, base + 2^bw - 1] are exceptions (bw 0 admits only the value 0:
// SafeUpperBound(base, 0) is 0, include/fls/utl/util.hpp:87-101); the size 128*bw + exceptions*sizeof(T) is
// minimised over settings with at most kExcLimit exceptions, ties going to the smaller base, then the larger bw.
// Without a better setting the vector is packed at full width with base 0 and no exceptions. The counts come from
// the sorted vector with one moving pointer per bit width. LL_PATCH then adds its compulsory links so that
// consecutive exceptions lie at most 2^bw apart (src/cor/prm/patch/ll_patch/ll_patch_compress.cpp:13-43).
template <typename T>
struct PatchPlan {
    unsigned              bw = 0;
    T                     base {};
    std::vector<uint16_t> pos;  // exception positions, ascending (LL: with its compulsory links)
};
template <typename T>
PatchPlan<T> plan_patch(const T* t, uint8_t var) {
    constexpr unsigned W = bits<T>();
    T                  s[1024];
    std::copy(t, t + 1024, s);
    std::sort(s, s + 1024);
    uint64_t best_size = 128ull * W;
    unsigned best_i = 0, best_bw = W;
    bool     found  = false;
    for (unsigned bw = 0; bw < W; ++bw) {
        const T  range = (T)((1ull << bw) - 1);
        unsigned j     = 0;  // first index above the window of base s[i]
        for (unsigned i = 0; i < 1024; ++i) {
            if (i && s[i] == s[i - 1]) continue;
            unsigned in_window;
            if (bw == 0) {
                in_window = 0;
                if (s[i] == 0)
                    while (in_window < 1024 && s[in_window] == 0) ++in_window;
            } else {
                const T upper = s[i] > (T)(~(T)0 - range) ? (T)~(T)0 : (T)(s[i] + range);
                if (j < i) j = i;
                while (j < 1024 && s[j] <= upper) ++j;
                in_window = j - i;
            }
            const uint64_t exc  = 1024 - in_window;
            const uint64_t size = 128ull * bw + exc * sizeof(T);
            if (exc > kExcLimit) continue;
            const bool better = size < best_size || (size == best_size && found && (i < best_i || (i == best_i && bw > best_bw)));
            if (better) best_size = size, best_i = i, best_bw = bw, found = true;
        }
    }
    PatchPlan<T> pp;
    if (!found) return pp.bw = W, pp.base = 0, pp;
    pp.bw   = best_bw;
    pp.base = s[best_i];
    const T lo = pp.base, hi = best_bw == 0 ? (T)0 : (lo > (T)(~(T)0 - (T)((1ull << best_bw) - 1)) ? (T)~(T)0 : (T)(lo + (T)((1ull << best_bw) - 1)));
    for (unsigned k = 0; k < 1024; ++k)
        if (t[k] < lo || t[k] > hi) pp.pos.push_back((uint16_t)k);
    if (var == kFforLL && !pp.pos.empty()) {
        const uint64_t        hop = 1ull << (best_bw < 11 ? best_bw : 11);
        std::vector<uint16_t> chain;
        uint64_t              cur = pp.pos[0];
        chain.push_back((uint16_t)cur);
        for (size_t i = 1; i < pp.pos.size(); ++i) {
            while (pp.pos[i] - cur > hop) cur += hop, chain.push_back((uint16_t)cur);
            cur = pp.pos[i];
            chain.push_back((uint16_t)cur);
        }
        pp.pos = std::move(chain);
    }
    return pp;
}

template <typename T>
void ffor(const T* in, uint8_t* out, unsigned bw, T base) {
    constexpr unsigned W = bits<T>(), L = 1024 / W;
    if (!bw) return;
    T* o = reinterpret_cast<T*>(out);
    std::memset(out, 0, 128ull * bw);
    const uint64_t m = bw >= 64 ? ~0ull : ((1ull << bw) - 1);
    for (unsigned i = 0; i < L; ++i)
        for (unsigned j = 0; j < W; ++j) {
            const uint64_t v   = (uint64_t)(T)(in[j * L + i] - base) & m;
            const unsigned bit = j * bw, w = bit / W, s = bit % W;
            o[w * L + i] |= (T)(v << s);
            if (s + bw > W) o[(w + 1) * L + i] |= (T)(v >> (W - s));
        }
}

struct Buf {
    std::vector<uint8_t> b;
    void put(const void* p, uint64_t n) { b.insert(b.end(), (const uint8_t*)p, (const uint8_t*)p + n); }
};

struct PageW {
    unsigned              arr_c = 0, vec_c = 0, rec_c = 0;  // rec_c: buffers stored as HZL chains
    Buf                   bufs[6], eps;
    uint32_t              ep[6] = {0};
    // Page::m_ttl_sz: header, offsets, entry points, the chains' record headers and the uncompressed buffers
    // (src/cor/lyt/page/page.cpp:44-56, :87-104)
    uint64_t              size() const {
        uint64_t s = 8 + 4ull * arr_c + eps.b.size() + 12ull * rec_c;
        for (unsigned i = 0; i < arr_c; ++i) s += bufs[i].b.size();
        return s;
    }
    void absorb(const std::vector<std::pair<const uint8_t*, uint64_t>>& parts) {
        if (vec_c) eps.put(ep, 4ull * arr_c);
        for (unsigned i = 0; i < arr_c; ++i) bufs[i].put(parts[i].first, parts[i].second), ep[i] += (uint32_t)parts[i].second;
        vec_c++;
    }
    void flush(uint8_t exp_t, uint8_t var, bool more, Buf& out) const {
        const uint8_t hdr[8] = {exp_t, var, 0, 0, (uint8_t)(vec_c & 0xff), (uint8_t)(vec_c >> 8), (uint8_t)more, 0};
        out.put(hdr, 8);
        uint32_t off = (uint32_t)(8 + 4 * arr_c + eps.b.size());
        for (unsigned i = 0; i < arr_c; ++i) out.put(&off, 4), off += (uint32_t)bufs[i].b.size();
        out.put(eps.b.data(), eps.b.size());
        for (unsigned i = 0; i < arr_c; ++i) out.put(bufs[i].b.data(), bufs[i].b.size());
    }
};

template <typename D>
void hzl_encode(const std::vector<D>& vals, Buf& out) {  // no exceptions; a full-width chunk ends the FFOR part
    uint64_t            pos = 0;
    std::vector<uint8_t> packed(8 * 1024);
    while (vals.size() - pos >= 1024) {
        D             base;
        const uint8_t bw = analyze<D>(vals.data() + pos, base);
        if (bw >= bits<D>()) break;
        uint8_t  hdr[12] = {bw, 0, 0, 0};
        uint64_t b64     = base;
        std::memcpy(hdr + 4, &b64, 8);
        out.put(hdr, 12);
        ffor<D>(vals.data() + pos, packed.data(), bw, base);
        out.put(packed.data(), 128ull * bw);
        pos += 1024;
    }
    const uint64_t rest    = (vals.size() - pos) * sizeof(D);
    uint8_t        hdr[12] = {255, 0, 0, 0};
    std::memcpy(hdr + 4, &rest, 8);
    out.put(hdr, 12);
    out.put(vals.data() + pos, rest);
}

using Parts = std::vector<std::pair<const uint8_t*, uint64_t>>;

// Byte buffers of one vector's parts: exceptions, positions, packed values, lane bases, RLE packed positions.
struct VecBufs {
    std::vector<uint8_t> exc = std::vector<uint8_t>(8192), pos = std::vector<uint8_t>(8192),
                         packed = std::vector<uint8_t>(8192), lbase = std::vector<uint8_t>(8192),
                         rle = std::vector<uint8_t>(8192);
    uint8_t  bw   = 0;
    uint64_t base = 0;
};

// One vector of an FFOR-shaped page from its transposed values t (modified): [lane bases] when the expression has
// RSUM (src/cor/prm/rsum/rsum_compress.cpp:10-22: the first 128 bytes, then unrsum in place), [exceptions]
// [positions] for a patch (PA_ANALYZE + S/B/LL_PATCH, s_patch_compress.cpp, b_patch_compress.cpp,
// ll_patch_compress.cpp), then [bw][base][packed] (ANALYZE / the patch plan + FFOR).
template <typename S>
void shaped_vector(const Expr& e, S* t, Parts& parts, VecBufs& vb, bool sgn) {
    unsigned o = 0;
    if (expr_rsum(e)) {
        S d[1024];
        std::memcpy(vb.lbase.data(), t, 128);
        parts[0] = {vb.lbase.data(), 128};
        unrsum<S>(t, d);
        std::copy(d, d + 1024, t);
        o = 1;
    }
    const uint8_t pv = expr_patch(e);
    if (!pv) {
        S base;
        vb.bw   = analyze<S>(t, base, sgn);
        vb.base = 0;
        std::memcpy(&vb.base, &base, sizeof(S));
        ffor<S>(t, vb.packed.data(), vb.bw, base);
    } else if constexpr (sizeof(S) >= 4) {
        const PatchPlan<S> pp = plan_patch<S>(t, pv);
        const uint16_t     n  = (uint16_t)pp.pos.size();
        for (uint16_t k = 0; k < n; ++k) std::memcpy(vb.exc.data() + k * sizeof(S), &t[pp.pos[k]], sizeof(S));
        parts[o] = {vb.exc.data(), (uint64_t)n * sizeof(S)};
        if (pv == kFforB) {
            std::memset(vb.pos.data(), 0, 128);
            for (uint16_t p : pp.pos) vb.pos[p >> 3] |= (uint8_t)(1u << (p & 7));
            parts[o + 1] = {vb.pos.data(), 128};
        } else if (pv == kFforS) {
            std::memcpy(vb.pos.data(), &n, 2);
            for (uint16_t k = 0; k < n; ++k) std::memcpy(vb.pos.data() + 2 + 2 * k, &pp.pos[k], 2);
            parts[o + 1] = {vb.pos.data(), 2ull + 2ull * n};
        } else {
            const uint16_t first = n ? pp.pos[0] : 0;
            std::memcpy(vb.pos.data(), &n, 2), std::memcpy(vb.pos.data() + 2, &first, 2);
            parts[o + 1] = {vb.pos.data(), 4};
        }
        // S and B leave the exceptions in place (FFOR keeps their low bw bits of value - base); LL stores in each
        // exception's slot the distance to the next one, minus one, plus base (the last slot's link is never
        // followed: it gets distance 1)
        if (pv == kFforLL)
            for (uint16_t k = 0; k < n; ++k) {
                const uint64_t gap = k + 1 < n ? (uint64_t)(pp.pos[k + 1] - pp.pos[k]) : 1;
                t[pp.pos[k]]       = (S)(pp.base + (S)(gap - 1));
            }
        vb.bw   = (uint8_t)pp.bw;
        vb.base = 0;
        std::memcpy(&vb.base, &pp.base, sizeof(S));
        ffor<S>(t, vb.packed.data(), pp.bw, pp.base);
        o += 2;
    }
    parts[o]     = {&vb.bw, 1};
    parts[o + 1] = {reinterpret_cast<const uint8_t*>(&vb.base), sizeof(S)};
    parts[o + 2] = {vb.packed.data(), 128ull * vb.bw};
}

// Whether consecutive values start a new run: bit inequality, or for doubles the reference's `!=` on double values
// (Vec::Init compares the column's own type, src/cor/lyt/vec.cpp:50-71): +0.0 and -0.0 stay in one run, and a NaN
// differs from everything, itself included.
template <typename T>
bool run_break(T a, T b, bool dbl) {
    if constexpr (sizeof(T) == 8) {
        if (dbl) {
            double x, y;
            std::memcpy(&x, &a, 8), std::memcpy(&y, &b, 8);
            return !(x == y);
        }
    }
    return a != b;
}

template <typename T>
int64_t encode_t(const Expr& e, const T* in, uint64_t n, Buf& out, uint64_t* meta, bool sgn, bool dbl = false) {
    using U = T;
    const uint64_t        n_vec = n / 1024;
    std::vector<uint32_t> codes;
    std::vector<uint64_t> dvals;
    if (e.dict) {  // dict_fac::perfect_create (src/cor/lyt/dic/dic_fac.cpp:259-330)
        codes.resize(n);
        std::unordered_map<U, uint32_t> first;
        std::vector<std::pair<U, uint64_t>> ents;  // value, count
        first.reserve(1024);
        for (uint64_t i = 0; i < n; ++i) {
            auto it = first.find(in[i]);
            if (it == first.end()) {
                it = first.emplace(in[i], (uint32_t)ents.size()).first;
                ents.push_back({in[i], 0});
            }
            ents[it->second].second++;
            codes[i] = it->second;
        }
        std::vector<uint32_t> order(ents.size());
        for (uint32_t k = 0; k < order.size(); ++k) order[k] = k;
        if (e.exp_t == EXP_DICT_FRQ)
            std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
                if (ents[a].second != ents[b].second) return ents[a].second > ents[b].second;
                return ents[a].first < ents[b].first;
            });
        std::vector<uint32_t> remap(order.size());
        dvals.resize(order.size());
        for (uint32_t c = 0; c < order.size(); ++c) remap[order[c]] = c, dvals[c] = (uint64_t)ents[order[c]].first;
        for (auto& c : codes) c = remap[c];
    }
    // 1) per-vector encodes, independent of each other: run on all host threads
    struct VecOut {
        std::vector<uint8_t>  bytes;
        std::vector<uint64_t> sizes;
    };
    std::vector<VecOut> vo(n_vec);
    auto encode_range = [&](uint64_t v0, uint64_t v1) {
        VecBufs  vb;
        T        t[1024];
        uint32_t ct[1024];
        for (uint64_t v = v0; v < v1; ++v) {
            const T* x = in + v * 1024;
            Parts    parts(e.arr_c);
            switch (e.exp_t) {
            case EXP_UNCOMPRESSED:
                transpose<T>(x, reinterpret_cast<T*>(vb.exc.data()));
                parts[0] = {vb.exc.data(), 1024ull * sizeof(T)};
                break;
            case EXP_FFOR:
            case EXP_DELTA:
                transpose<T>(x, t);
                shaped_vector<T>(e, t, parts, vb, sgn);
                break;
            case EXP_RLE:
            case EXP_DICT_RLE: {  // DICT_RLE: Vec::Init's split of the codes (src/cor/lyt/vec.cpp:72-96) first
                const uint32_t* c   = e.exp_t == EXP_DICT_RLE ? codes.data() + v * 1024 : nullptr;
                const unsigned  rsz = c ? 4 : sizeof(T);
                uint16_t        pos[1024], posT[1024], dl[1024];
                uint64_t        r = 0;
                for (unsigned i = 0; i < 1024; ++i) {
                    const bool brk = i && (c ? c[i] != c[i - 1] : run_break<T>(x[i], x[i - 1], dbl));
                    if (i == 0 || brk) {
                        if (i) r++;
                        std::memcpy(vb.exc.data() + r * rsz, c ? (const void*)&c[i] : (const void*)&x[i], rsz);
                    }
                    pos[i] = (uint16_t)r;
                }
                parts[0] = {vb.exc.data(), (r + 1) * rsz};
                transpose<uint16_t>(pos, posT);
                if (e.var == 1 || e.var == kRleRec) {
                    std::memcpy(vb.pos.data(), posT, 2048);
                    parts[1] = {vb.pos.data(), 2048};
                } else {
                    std::memcpy(vb.pos.data(), posT, 128);
                    parts[1] = {vb.pos.data(), 128};
                    unrsum<uint16_t>(posT, dl);
                    uint16_t base;
                    vb.bw   = analyze<uint16_t>(dl, base);
                    vb.base = base;
                    ffor<uint16_t>(dl, vb.rle.data(), vb.bw, base);
                    parts[2] = {&vb.bw, 1};
                    parts[3] = {reinterpret_cast<const uint8_t*>(&vb.base), 2};
                    parts[4] = {vb.rle.data(), 128ull * vb.bw};
                }
                break;
            }
            default:  // dictionary codes (u32)
                transpose<uint32_t>(codes.data() + v * 1024, ct);
                if (e.var == 1) {
                    std::memcpy(vb.exc.data(), ct, 4096);
                    parts[0] = {vb.exc.data(), 4096};
                } else {
                    shaped_vector<uint32_t>(e, ct, parts, vb, false);
                }
            }
            VecOut& o = vo[v];
            for (auto& pp : parts) o.bytes.insert(o.bytes.end(), pp.first, pp.first + pp.second), o.sizes.push_back(pp.second);
        }
    };
    {
        const unsigned hw = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
        const unsigned nt = n_vec < 64 ? 1 : hw;
        std::vector<std::thread> th;
        for (unsigned k = 0; k < nt; ++k) th.emplace_back(encode_range, n_vec * k / nt, n_vec * (k + 1) / nt);
        for (auto& x : th) x.join();
    }
    // 2) sequential page assembly (Page::Absorb order)
    const uint64_t page_c = n_vec ? (n_vec + kMaxVecPerPage - 1) / kMaxVecPerPage : 1;
    const uint64_t per_pg = n_vec ? (n_vec + page_c - 1) / page_c : 0;
    uint64_t       v      = 0;
    constexpr unsigned W  = bits<T>();
    bool           first_page = true;
    for (;;) {
        PageW pg;
        pg.arr_c = e.arr_c;
        pg.rec_c = (unsigned)__builtin_popcount(expr_rec(e));
        const uint64_t stop = std::min(n_vec, v + per_pg);
        for (; v < stop; ++v) {
            std::vector<std::pair<const uint8_t*, uint64_t>> parts(e.arr_c);
            uint64_t off = 0, add = 4ull * e.arr_c;
            for (unsigned i = 0; i < e.arr_c; ++i) parts[i] = {vo[v].bytes.data() + off, vo[v].sizes[i]}, off += vo[v].sizes[i], add += vo[v].sizes[i];
            if (pg.vec_c && pg.size() + add >= kDataPageMax) break;  // TryFlushVec limit (compressor.cpp:47)
            pg.absorb(parts);
            std::vector<uint8_t>().swap(vo[v].bytes);
        }
        const bool more = v < n_vec;
        if (e.dict) {
            const uint8_t  hdr[4] = {EXP_DICT, 0, 0, 0};
            const uint32_t ne     = (uint32_t)dvals.size();
            out.put(hdr, 4), out.put(&ne, 4);
            if (W <= 32) {
                std::vector<uint32_t> dv(dvals.size());
                for (size_t k = 0; k < dv.size(); ++k) dv[k] = (uint32_t)dvals[k];
                hzl_encode<uint32_t>(dv, out);
            } else {
                hzl_encode<uint64_t>(dvals, out);
            }
        }
        // buffers stored as one HZL chain per page (THzlCmp::compress_buf + finalize_buf, hzl_cmp.cpp:43-176)
        const unsigned rec = expr_rec(e);
        for (unsigned b = 0; b < 2 && pg.vec_c; ++b) {
            if (!(rec >> b & 1)) continue;
            Buf h;
            if (expr_rec_width(e) == 32) {
                std::vector<uint32_t> xv(pg.bufs[b].b.size() / 4);
                if (!xv.empty()) std::memcpy(xv.data(), pg.bufs[b].b.data(), xv.size() * 4);
                hzl_encode<uint32_t>(xv, h);
            } else {
                std::vector<uint64_t> xv(pg.bufs[b].b.size() / 8);
                if (!xv.empty()) std::memcpy(xv.data(), pg.bufs[b].b.data(), xv.size() * 8);
                hzl_encode<uint64_t>(xv, h);
            }
            pg.bufs[b].b.swap(h.b);
        }
        if (first_page && meta) *meta = out.b.size() + 8 + 4ull * e.arr_c * std::max(1u, pg.vec_c);
        first_page = false;
        pg.flush(e.exp_t, e.var, more, out);
        if (!more) break;
    }
    (void)W;
    return (int64_t)out.b.size();
}

}  // namespace

int64_t encode_column(uint8_t dtype, uint8_t expr_id, const void* in, uint64_t n, uint8_t* out, uint64_t cap,
                      uint64_t* metadata_size) {
    Expr e;
    if (!expr_from_id(dtype, expr_id, e) || dtype == DT_STR) return -1;
    if (!expr_fits_type(e, dtype)) return -1;
    if (n % 1024) return -1;
    Buf     b;
    int64_t r = -1;
    switch (dtype_bits(dtype)) {
    case 8: r = encode_t<uint8_t>(e, (const uint8_t*)in, n, b, metadata_size, false); break;
    case 16: r = encode_t<uint16_t>(e, (const uint16_t*)in, n, b, metadata_size, false); break;
    case 32: r = encode_t<uint32_t>(e, (const uint32_t*)in, n, b, metadata_size, false); break;
    case 64: r = encode_t<uint64_t>(e, (const uint64_t*)in, n, b, metadata_size, dtype == DT_I64, dtype == DT_DOUBLE); break;
    default: return -1;
    }
    if (r < 0) return r;
    if ((uint64_t)r > cap) return -2;
    std::memcpy(out, b.b.data(), (size_t)r);
    return r;
}

}  // namespace fls
