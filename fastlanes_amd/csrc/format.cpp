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

bool expr_from_id(uint8_t dtype, uint8_t id, Expr& e) {
    if (dtype == DT_STR) {
        if (id == 13) return e = {EXP_FSST, 1, 2, true}, true;
        return false;
    }
    switch (id) {
    case 0: e = {EXP_UNCOMPRESSED, 1, 1, false}; return true;
    case 1: e = {EXP_FFOR, 1, 3, false}; return true;
    case 2: e = {EXP_FFOR, kFforS, 5, false}; return dtype_bits(dtype) >= 32 && dtype != DT_DOUBLE;   // S_PATCH
    case 3: e = {EXP_FFOR, kFforSRec, 5, false}; return dtype_bits(dtype) == 64 && dtype != DT_DOUBLE;  // S_PATCH_REC
    case 4: e = {EXP_FFOR, kFforB, 5, false}; return dtype_bits(dtype) >= 32 && dtype != DT_DOUBLE;   // B_PATCH
    case 5: e = {EXP_FFOR, kFforLL, 5, false}; return dtype_bits(dtype) >= 32 && dtype != DT_DOUBLE;  // LL_PATCH
    case 6: e = {EXP_RLE, 2, 5, false}; return true;
    case 7: e = {EXP_RLE, 1, 2, false}; return true;
    case 11: e = {EXP_DELTA, 1, 4, false}; return true;
    case 16: e = {EXP_DICT, 1, 1, true}; return true;
    case 17: e = {EXP_DICT, 2, 3, true}; return true;
    case 22: e = {EXP_DICT, 7, 4, true}; return true;
    case 27: e = {EXP_DICT_RLE, kDictRleRec, 5, true}; return dtype != DT_DOUBLE;
    case 28: e = {EXP_DICT_FRQ, 1, 1, true}; return true;
    case 29: e = {EXP_DICT_FRQ, 2, 3, true}; return true;
    default: return false;
    }
}

bool expr_from_page(uint8_t exp_t, uint8_t var, Expr& e) {
    switch (exp_t) {
    case EXP_UNCOMPRESSED: e = {exp_t, var, 1, false}; return var == 1;
    case EXP_FFOR:
        if (var == 1) return e = {exp_t, var, 3, false}, true;
        if (var == kFforLL || var == kFforS || var == kFforB || var == kFforSRec) return e = {exp_t, var, 5, false}, true;
        return false;
    case EXP_DELTA: e = {exp_t, var, 4, false}; return var == 1;
    case EXP_RLE:
        if (var == 1) return e = {exp_t, var, 2, false}, true;
        if (var == 2) return e = {exp_t, var, 5, false}, true;
        return false;
    case EXP_DICT:
    case EXP_DICT_FRQ:
    case EXP_DICT_VAL:
        if (var == 1) return e = {exp_t, var, 1, true}, true;
        if (var == 2) return e = {exp_t, var, 3, true}, true;
        if (var == 7) return e = {exp_t, var, 4, true}, true;
        return false;
    case EXP_DICT_RLE: e = {exp_t, var, 5, true}; return var == kDictRleRec;
    case EXP_FSST: e = {exp_t, var, 2, true}; return var == 1;
    default: return false;
    }
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
    if (dtype == DT_DOUBLE && e.exp_t != EXP_UNCOMPRESSED) return "double columns support UNCOMPRESSED only";
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
    // DICT_RLE / S_PATCH_REC: buffer 0 is an HZL chain (u32 run codes / u64 exception values) whose entry points
    // address the decoded values
    const bool     srec = e.exp_t == EXP_FFOR && e.var == kFforSRec;
    const bool     drle = e.exp_t == EXP_DICT_RLE || srec;
    const unsigned rdw  = srec ? 64 : 32;
    if (srec && W != 64) return "S_PATCH_REC needs 64-bit values";
    if (drle) {
        uint64_t used = 0;
        auto     err  = parse_hzl(r, off + boff[0], rdw, ~0u, pe.run_hzl, used);
        if (!err.empty()) return "run codes: " + err;
        if (used != boff[1] - boff[0]) return "run codes do not fill their buffer";
        for (const auto& c : pe.run_hzl) pe.run_codes += c.n;
        if (rdw / 8 * pe.run_codes > 0xffffffffull) return "recursive buffer exceeds 32-bit entry points";
    }
    auto cap = [&](unsigned b) {
        return drle && b == 0 ? rdw / 8 * pe.run_codes : (b + 1 < arr_c ? boff[b + 1] : avail) - boff[b];
    };
    auto           need = [&](unsigned v, unsigned b, uint64_t n) -> bool {
        const uint64_t s = ep(v, b);
        const uint64_t lim = v + 1 < pe.vec_c ? ep(v + 1, b) : cap(b);
        return s <= lim && lim <= cap(b) && n <= lim - s;
    };
    auto bwc = [&](unsigned v, unsigned b) { return (unsigned)pg[boff[b] + ep(v, b)]; };
    uint64_t end = 0;  // bytes of the last buffer used by the last vector
    for (unsigned v = 0; v < pe.vec_c; ++v) {
        bool ok = true;
        switch (e.exp_t) {
        case EXP_UNCOMPRESSED: ok = need(v, 0, 1024ull * sz); break;
        case EXP_FFOR:
            if (e.var != 1) {  // patched: [exc values][exc positions | bitmap | count+first][bw][base][packed]
                if (W < 32) return "patched FFOR needs 32- or 64-bit values";
                ok = need(v, 2, 1) && need(v, 3, sz);
                if (!ok) break;
                const unsigned bw = bwc(v, 2);
                ok = bw <= W && need(v, 4, 128ull * bw);
                if (!ok) break;
                const uint8_t* b1  = pg + boff[1] + ep(v, 1);
                uint32_t       exc = 0;
                const uint8_t  pv  = srec ? kFforS : e.var;
                if (pv == kFforB) {
                    ok = need(v, 1, 128);
                    if (!ok) break;
                    for (unsigned k = 0; k < 128; ++k) exc += (uint32_t)__builtin_popcount(b1[k]);
                } else {
                    ok = need(v, 1, pv == kFforS ? 2 : 4);
                    if (!ok) break;
                    exc = (uint32_t)b1[0] | ((uint32_t)b1[1] << 8);
                    if (pv == kFforS) {
                        ok = exc <= 1024 && need(v, 1, 2 + 2ull * exc);
                        for (uint32_t k = 0; ok && k < exc; ++k) ok = ((uint32_t)b1[2 + 2 * k] | ((uint32_t)b1[3 + 2 * k] << 8)) < 1024;
                    } else {
                        ok = exc <= 1024 && (exc == 0 || ((uint32_t)b1[2] | ((uint32_t)b1[3] << 8)) < 1024);
                    }
                    if (!ok) break;
                }
                ok = need(v, 0, (uint64_t)exc * sz);
                break;
            }
            [[fallthrough]];
        case EXP_DELTA: {
            const unsigned o = e.exp_t == EXP_DELTA;
            if (o) ok &= need(v, 0, 128);
            ok = ok && need(v, o, 1) && need(v, o + 1, sz);
            if (!ok) break;
            const unsigned bw = bwc(v, o);
            ok = bw <= W && need(v, o + 2, 128ull * bw);
            break;
        }
        case EXP_RLE: {
            ok = need(v, 0, sz);  // at least one run value
            if (e.var == 1) ok = ok && need(v, 1, 2048);
            else {
                ok = ok && need(v, 1, 128) && need(v, 2, 1) && need(v, 3, 2);
                if (!ok) break;
                const unsigned bw = bwc(v, 2);
                ok = bw <= 16 && need(v, 4, 128ull * bw);
            }
            break;
        }
        case EXP_FSST: ok = need(v, 1, 4096) && need(v, 0, 0); break;
        case EXP_DICT_RLE: {  // RLE DELTA_NO_PATCH positions over u32 run codes
            ok = ep(v, 0) % 4 == 0 && need(v, 0, 4) && need(v, 1, 128) && need(v, 2, 1) && need(v, 3, 2);
            if (!ok) break;
            const unsigned bw = bwc(v, 2);
            ok = bw <= 16 && need(v, 4, 128ull * bw);
            break;
        }
        default: {  // dictionary codes
            if (e.var == 1) ok = need(v, 0, 4096);
            else {
                const unsigned o = e.var == 7;
                if (o) ok &= need(v, 0, 128);
                ok = ok && need(v, o, 1) && need(v, o + 1, 4);
                if (!ok) break;
                const unsigned bw = bwc(v, o);
                ok = bw <= 32 && need(v, o + 2, 128ull * bw);
            }
        }
        }
        if (!ok) return "vector " + std::to_string(v) + " exceeds its page buffers";
    }
    for (unsigned b = 0; b < arr_c; ++b) pe.boff[b] = (uint32_t)boff[b];
    {  // fixed per-vector sizes of the metadata buffers (0 = variable size)
        unsigned fsz[8] = {};
        switch (e.exp_t) {
        case EXP_UNCOMPRESSED: fsz[0] = 1024 * sz; break;
        case EXP_FFOR:
            if (e.var == 1) fsz[0] = 1, fsz[1] = sz;
            else fsz[2] = 1, fsz[3] = sz, fsz[1] = e.var == kFforB ? 128 : e.var == kFforLL ? 4 : 0;  // S, S_REC: 0
            break;
        case EXP_DELTA: fsz[0] = 128, fsz[1] = 1, fsz[2] = sz; break;
        case EXP_RLE:
            if (e.var == 1) fsz[1] = 2048;
            else fsz[1] = 128, fsz[2] = 1, fsz[3] = 2;
            break;
        case EXP_FSST: fsz[1] = 4096; break;
        case EXP_DICT_RLE: fsz[1] = 128, fsz[2] = 1, fsz[3] = 2; break;
        default:
            if (e.var == 1) fsz[0] = 4096;
            else if (e.var == 2) fsz[0] = 1, fsz[1] = 4;
            else fsz[0] = 128, fsz[1] = 1, fsz[2] = 4;
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
        switch (e.exp_t) {
        case EXP_UNCOMPRESSED: tail = 1024ull * sz; break;
        case EXP_FFOR:
        case EXP_DELTA: tail = 128ull * bwc(v, b - 2); break;  // patched FFOR: b - 2 is its bw buffer too
        case EXP_RLE: tail = e.var == 1 ? 2048 : 128ull * bwc(v, b - 2); break;
        case EXP_FSST: tail = 4096; break;
        default: tail = e.var == 1 ? 4096 : 128ull * bwc(v, b - 2); break;
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
template <typename T>
uint8_t analyze(const T* in, T& base) {  // base = min, bw = exact bit length of max - min
    T mn = in[0], mx = in[0];
    for (unsigned i = 1; i < 1024; ++i) mn = in[i] < mn ? in[i] : mn, mx = in[i] > mx ? in[i] : mx;
    base             = mn;
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
    unsigned              arr_c = 0, vec_c = 0;
    Buf                   bufs[6], eps;
    uint32_t              ep[6] = {0};
    uint64_t              size() const {
        uint64_t s = 8 + 4ull * arr_c + eps.b.size();
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

template <typename T>
int64_t encode_t(const Expr& e, const T* in, uint64_t n, Buf& out, uint64_t* meta) {
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
        std::vector<uint8_t> b0(8192), b1(8192), b2(8192), b3(8192), b4(8192);
        T                    t[1024], d[1024];
        uint32_t             ct[1024], cd[1024];
        uint8_t              bwb;
        uint64_t             baseb;
        for (uint64_t v = v0; v < v1; ++v) {
            const T* x = in + v * 1024;
            std::vector<std::pair<const uint8_t*, uint64_t>> parts(e.arr_c);
            auto ffor_parts = [&](auto* src, unsigned o, uint8_t* packed) {
                using S = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
                S             base;
                const uint8_t bw = analyze<S>(src, base);
                bwb              = bw;
                baseb            = 0;
                std::memcpy(&baseb, &base, sizeof(S));
                ffor<S>(src, packed, bw, base);
                parts[o]     = {&bwb, 1};
                parts[o + 1] = {reinterpret_cast<const uint8_t*>(&baseb), sizeof(S)};
                parts[o + 2] = {packed, 128ull * bw};
            };
            switch (e.exp_t) {
            case EXP_UNCOMPRESSED:
                transpose<T>(x, reinterpret_cast<T*>(b0.data()));
                parts[0] = {b0.data(), 1024ull * sizeof(T)};
                break;
            case EXP_FFOR:
                transpose<T>(x, t);
                if (e.var == 1) {
                    ffor_parts(t, 0, b2.data());
                } else if constexpr (sizeof(T) >= 4) {
                    const uint8_t      pv = e.var == kFforSRec ? kFforS : e.var;  // S_PATCH_REC vectors are S_PATCH
                    const PatchPlan<T> pp = plan_patch<T>(t, pv);
                    const uint16_t     n  = (uint16_t)pp.pos.size();
                    for (uint16_t k = 0; k < n; ++k) std::memcpy(b0.data() + k * sizeof(T), &t[pp.pos[k]], sizeof(T));
                    parts[0] = {b0.data(), (uint64_t)n * sizeof(T)};
                    if (pv == kFforB) {
                        std::memset(b1.data(), 0, 128);
                        for (uint16_t p : pp.pos) b1[p >> 3] |= (uint8_t)(1u << (p & 7));
                        parts[1] = {b1.data(), 128};
                    } else if (pv == kFforS) {
                        std::memcpy(b1.data(), &n, 2);
                        for (uint16_t k = 0; k < n; ++k) std::memcpy(b1.data() + 2 + 2 * k, &pp.pos[k], 2);
                        parts[1] = {b1.data(), 2ull + 2ull * n};
                    } else {
                        const uint16_t first = n ? pp.pos[0] : 0;
                        std::memcpy(b1.data(), &n, 2), std::memcpy(b1.data() + 2, &first, 2);
                        parts[1] = {b1.data(), 4};
                    }
                    // S and B leave the exceptions in place (FFOR keeps their low bw bits of value - base); LL
                    // stores in each exception's slot the distance to the next one, minus one, plus base (the last
                    // slot's link is never followed: it gets distance 1)
                    if (pv == kFforLL)
                        for (uint16_t k = 0; k < n; ++k) {
                            const uint64_t gap = k + 1 < n ? (uint64_t)(pp.pos[k + 1] - pp.pos[k]) : 1;
                            t[pp.pos[k]]       = (T)(pp.base + (T)(gap - 1));
                        }
                    bwb   = (uint8_t)pp.bw;
                    baseb = 0;
                    std::memcpy(&baseb, &pp.base, sizeof(T));
                    ffor<T>(t, b2.data(), pp.bw, pp.base);
                    parts[2] = {&bwb, 1};
                    parts[3] = {reinterpret_cast<const uint8_t*>(&baseb), sizeof(T)};
                    parts[4] = {b2.data(), 128ull * pp.bw};
                }
                break;
            case EXP_DELTA:
                transpose<T>(x, t);
                std::memcpy(b0.data(), t, 128);
                parts[0] = {b0.data(), 128};
                unrsum<T>(t, d);
                ffor_parts(d, 1, b3.data());
                break;
            case EXP_RLE: {
                uint16_t pos[1024], posT[1024], dl[1024];
                uint64_t r = 0;
                for (unsigned i = 0; i < 1024; ++i) {
                    if (i == 0 || x[i] != x[i - 1]) {
                        if (i) r++;
                        std::memcpy(b0.data() + r * sizeof(T), &x[i], sizeof(T));
                    }
                    pos[i] = (uint16_t)r;
                }
                parts[0] = {b0.data(), (r + 1) * sizeof(T)};
                transpose<uint16_t>(pos, posT);
                if (e.var == 1) {
                    std::memcpy(b1.data(), posT, 2048);
                    parts[1] = {b1.data(), 2048};
                } else {
                    std::memcpy(b1.data(), posT, 128);
                    parts[1] = {b1.data(), 128};
                    unrsum<uint16_t>(posT, dl);
                    ffor_parts(dl, 2, b4.data());
                }
                break;
            }
            case EXP_DICT_RLE: {  // Vec::Init split of the codes (src/cor/lyt/vec.cpp:72-96), then RLE DELTA_NO_PATCH
                const uint32_t* c = codes.data() + v * 1024;
                uint16_t        pos[1024], posT[1024], dl[1024];
                uint64_t        r = 0;
                for (unsigned i = 0; i < 1024; ++i) {
                    if (i == 0 || c[i] != c[i - 1]) {
                        if (i) r++;
                        std::memcpy(b0.data() + r * 4, &c[i], 4);
                    }
                    pos[i] = (uint16_t)r;
                }
                parts[0] = {b0.data(), (r + 1) * 4};
                transpose<uint16_t>(pos, posT);
                std::memcpy(b1.data(), posT, 128);
                parts[1] = {b1.data(), 128};
                unrsum<uint16_t>(posT, dl);
                ffor_parts(dl, 2, b4.data());
                break;
            }
            default: {
                transpose<uint32_t>(codes.data() + v * 1024, ct);
                if (e.var == 1) {
                    std::memcpy(b0.data(), ct, 4096);
                    parts[0] = {b0.data(), 4096};
                } else if (e.var == 2) {
                    ffor_parts(ct, 0, b2.data());
                } else {
                    std::memcpy(b0.data(), ct, 128);
                    parts[0] = {b0.data(), 128};
                    unrsum<uint32_t>(ct, cd);
                    ffor_parts(cd, 1, b3.data());
                }
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
        if (e.exp_t == EXP_DICT_RLE && pg.vec_c) {  // the page's run codes as one HZL chain (THzlCmp<idx_t>, hzl_cmp.cpp:43-176)
            std::vector<uint32_t> rc(pg.bufs[0].b.size() / 4);
            if (!rc.empty()) std::memcpy(rc.data(), pg.bufs[0].b.data(), rc.size() * 4);
            Buf h;
            hzl_encode<uint32_t>(rc, h);
            pg.bufs[0].b.swap(h.b);
        }
        if (e.exp_t == EXP_FFOR && e.var == kFforSRec && pg.vec_c) {  // exception values, THzlCmp<uint64_t>
            std::vector<uint64_t> xv(pg.bufs[0].b.size() / 8);
            if (!xv.empty()) std::memcpy(xv.data(), pg.bufs[0].b.data(), xv.size() * 8);
            Buf h;
            hzl_encode<uint64_t>(xv, h);
            pg.bufs[0].b.swap(h.b);
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
    if (dtype == DT_DOUBLE && e.exp_t != EXP_UNCOMPRESSED) return -1;
    if (n % 1024) return -1;
    Buf     b;
    int64_t r = -1;
    switch (dtype_bits(dtype)) {
    case 8: r = encode_t<uint8_t>(e, (const uint8_t*)in, n, b, metadata_size); break;
    case 16: r = encode_t<uint16_t>(e, (const uint16_t*)in, n, b, metadata_size); break;
    case 32: r = encode_t<uint32_t>(e, (const uint32_t*)in, n, b, metadata_size); break;
    case 64: r = encode_t<uint64_t>(e, (const uint64_t*)in, n, b, metadata_size); break;
    default: return -1;
    }
    if (r < 0) return r;
    if ((uint64_t)r > cap) return -2;
    std::memcpy(out, b.b.data(), (size_t)r);
    return r;
}

}  // namespace fls
