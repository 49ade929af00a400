// fsst_encode.cpp -- host encoder for FSST string columns.
//
// Symbol table: the iterative candidate-gain construction of Boncz, Neumann, Leis, "FSST: Fast Random Access String
// Compression", PVLDB 13(11) 2020, Sec. 4 (five rounds over a ~32 KB sample; candidates are the symbols found by a
// greedy longest-match parse and the concatenations of adjacent symbols; gain = count * length; keep the best 255).
// Serialized as fsst_export writes it (reference src/cor/prm/fsst/libfsst.cpp:582-615): u64 version, zero-terminated
// flag, lenHisto[8], symbol bytes ordered by length 2..8 then 1 (the code order fsst_import assigns, :617-652).
// Column layout per vector (reference src/cor/lyt/vec.cpp:160-207, include/fls/cor/exp/fsst/fsst_exp.hpp:35-52):
// buffer 0 = compressed strings back to back, buffer 1 = u32 cumulative end offsets, transposed.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "format.hpp"

namespace fls {
namespace {

struct Sym {
    uint64_t val;
    uint8_t  len;
};

struct Table {
    Sym      s[255];
    unsigned n = 0;
    int16_t  one[256];
    std::vector<uint32_t> start;  // 65537 bucket starts over the first two bytes
    std::vector<uint8_t>  codes;  // codes of length >= 2, longest first within a bucket

    void index() {
        std::vector<uint32_t> cnt(65536, 0);
        for (auto& o : one) o = -1;
        for (unsigned c = 0; c < n; ++c) {
            if (s[c].len == 1) one[s[c].val & 0xff] = (int16_t)c;
            else cnt[s[c].val & 0xffff]++;
        }
        start.assign(65537, 0);
        for (unsigned k = 0; k < 65536; ++k) start[k + 1] = start[k] + cnt[k];
        codes.assign(start[65536] + 1, 0);
        std::vector<uint32_t> fill(start.begin(), start.end() - 1);
        for (unsigned l = 8; l >= 2; --l)
            for (unsigned c = 0; c < n; ++c)
                if (s[c].len == l) codes[fill[s[c].val & 0xffff]++] = (uint8_t)c;
    }
    int match(const uint8_t* p, size_t rem) const {
        if (rem >= 2) {
            const unsigned k = p[0] | ((unsigned)p[1] << 8);
            for (uint32_t q = start[k]; q < start[k + 1]; ++q) {
                const unsigned c = codes[q], l = s[c].len;
                if (l <= rem && std::memcmp(&s[c].val, p, l) == 0) return (int)c;
            }
        }
        return one[p[0]];
    }
    size_t compress(const uint8_t* in, size_t len, uint8_t* out) const {
        size_t i = 0, o = 0;
        while (i < len) {
            const int c = match(in + i, len - i);
            if (c < 0) out[o++] = 255, out[o++] = in[i++];
            else out[o++] = (uint8_t)c, i += s[c].len;
        }
        return o;
    }
};

uint64_t mix(uint64_t x) {
    x ^= x >> 33, x *= 0xff51afd7ed558ccdULL, x ^= x >> 33, x *= 0xc4ceb9fe1a85ec53ULL, x ^= x >> 33;
    return x;
}

struct Cand {
    uint64_t val;
    uint8_t  len;
    uint64_t cnt;
};

void build_table(const uint32_t* offsets, const uint8_t* bytes, uint64_t n, Table& t) {
    const uint64_t total  = offsets[n] - offsets[0];
    const uint64_t stride = total > (1u << 15) ? total / (1u << 15) + 1 : 1;
    const uint64_t cap    = 1 << 16;
    std::vector<Cand> tab(cap);
    auto add = [&](uint64_t v, uint8_t l) {
        uint64_t h = mix(v * 31 + l) & (cap - 1);
        for (;;) {
            if (tab[h].len == 0) return void(tab[h] = Cand {v, l, 1});
            if (tab[h].len == l && tab[h].val == v) return void(tab[h].cnt++);
            h = (h + 1) & (cap - 1);
        }
    };
    for (int round = 0; round < 5; ++round) {
        t.index();
        std::fill(tab.begin(), tab.end(), Cand {0, 0, 0});
        for (uint64_t si = 0; si < n; si += stride) {
            const uint8_t* p   = bytes + offsets[si];
            const size_t   rem = offsets[si + 1] - offsets[si];
            size_t         i   = 0;
            bool           has_prev = false;
            uint64_t       pval = 0;
            uint8_t        plen = 0;
            while (i < rem) {
                const int     c = t.match(p + i, rem - i);
                const uint8_t l = c >= 0 ? t.s[c].len : 1;
                uint64_t      v = 0;
                std::memcpy(&v, p + i, l);
                add(v, l);
                if (l != 1) add(p[i], 1);
                if (has_prev && plen + l <= 8) add(pval | (v << (8 * plen)), (uint8_t)(plen + l));
                has_prev = true, pval = v, plen = l;
                i += l;
            }
        }
        std::vector<Cand> c;
        for (auto& x : tab)
            if (x.len) c.push_back(x);
        std::sort(c.begin(), c.end(), [](const Cand& a, const Cand& b) {
            const uint64_t ga = a.cnt * a.len, gb = b.cnt * b.len;
            if (ga != gb) return ga > gb;
            if (a.len != b.len) return a.len > b.len;
            return a.val < b.val;
        });
        t.n = 0;
        for (auto& x : c) {
            if (t.n == 255) break;
            t.s[t.n++] = Sym {x.val, x.len};
        }
    }
    // code order of fsst_import: lengths 2..8, then 1
    Sym      sorted[255];
    unsigned k = 0;
    for (unsigned L = 2; L <= 9; ++L)
        for (unsigned c = 0; c < t.n; ++c)
            if (t.s[c].len == (L == 9 ? 1u : L)) sorted[k++] = t.s[c];
    std::memcpy(t.s, sorted, sizeof(Sym) * k);
    t.n = k;
    t.index();
}

std::vector<uint8_t> export_table(const Table& t) {
    unsigned lh[8] = {0};
    for (unsigned c = 0; c < t.n; ++c) lh[t.s[c].len - 1]++;
    std::vector<uint8_t> b(17);
    const uint64_t       version = (20190218ull << 32) | ((uint64_t)lh[1] << 24) | ((uint64_t)t.n << 8) | 1;
    std::memcpy(b.data(), &version, 8);
    b[8] = 0;
    for (int i = 0; i < 8; ++i) b[9 + i] = (uint8_t)lh[i];
    for (unsigned c = 0; c < t.n; ++c)
        for (unsigned j = 0; j < t.s[c].len; ++j) b.push_back(((const uint8_t*)&t.s[c].val)[j]);
    return b;
}

constexpr unsigned kOrder8[8] = {0, 4, 2, 6, 1, 5, 3, 7};
inline unsigned orig_of(unsigned k) { return (k >> 7) + 8 * kOrder8[(k >> 4) & 7] + 64 * (k & 15); }

}  // namespace

int64_t encode_str_column(uint8_t expr_id, const uint32_t* offsets, const uint8_t* bytes, uint64_t n, uint8_t* out,
                          uint64_t cap) {
    if (expr_id != 13 || n % 1024) return -1;
    Table t;
    build_table(offsets, bytes, n, t);
    const std::vector<uint8_t> table = export_table(t);
    const uint64_t             n_vec = n / 1024;
    // per-vector compression on all host threads
    std::vector<std::vector<uint8_t>> comp(n_vec);
    std::vector<uint32_t>             endsT(n_vec * 1024);
    auto run = [&](uint64_t v0, uint64_t v1) {
        std::vector<uint8_t> buf;
        for (uint64_t v = v0; v < v1; ++v) {
            uint64_t raw = offsets[(v + 1) * 1024] - offsets[v * 1024];
            buf.resize(2 * raw + 16);
            uint32_t c = 0, ends[1024];
            for (unsigned i = 0; i < 1024; ++i) {
                const uint64_t s = v * 1024 + i;
                c += (uint32_t)t.compress(bytes + offsets[s], offsets[s + 1] - offsets[s], buf.data() + c);
                ends[i] = c;
            }
            comp[v].assign(buf.begin(), buf.begin() + c);
            for (unsigned k = 0; k < 1024; ++k) endsT[v * 1024 + k] = ends[orig_of(k)];
        }
    };
    {
        const unsigned hw = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
        const unsigned nt = n_vec < 64 ? 1 : hw;
        std::vector<std::thread> th;
        for (unsigned k = 0; k < nt; ++k) th.emplace_back(run, n_vec * k / nt, n_vec * (k + 1) / nt);
        for (auto& x : th) x.join();
    }
    // page entries: [dict_hdr_t{FSST}][table][data page]
    std::vector<uint8_t> o;
    auto put = [&](const void* p, uint64_t k) { o.insert(o.end(), (const uint8_t*)p, (const uint8_t*)p + k); };
    const uint64_t page_c = n_vec ? (n_vec + kMaxVecPerPage - 1) / kMaxVecPerPage : 1;
    const uint64_t per_pg = n_vec ? (n_vec + page_c - 1) / page_c : 0;
    uint64_t       v      = 0;
    if (n_vec == 0) {  // an empty column is one empty page entry
        const uint8_t dh[8] = {EXP_FSST, 0, 0, 0, 0, 0, 0, 0}, hdr[8] = {EXP_FSST, 1, 0, 0, 0, 0, 0, 0};
        const uint32_t offs[2] = {16, 16};
        put(dh, 8), put(table.data(), table.size()), put(hdr, 8), put(offs, 8);
    }
    while (n_vec) {
        const uint64_t stop = std::min(n_vec, v + per_pg);
        // a page must stay below the reference's data-page limit
        uint64_t bytes0 = 0, stop2 = v;
        while (stop2 < stop && 8 + 8 + 8 * (stop2 - v) + bytes0 + comp[stop2].size() + 4096 * (stop2 - v + 1) < kDataPageMax)
            bytes0 += comp[stop2].size(), ++stop2;
        if (stop2 == v) stop2 = v + 1;
        const uint32_t vec_c = (uint32_t)(stop2 - v);
        const bool     more  = stop2 < n_vec;
        const uint8_t  dh[8] = {EXP_FSST, 0, 0, 0, 0, 0, 0, 0};
        put(dh, 8);
        put(table.data(), table.size());
        const uint8_t hdr[8] = {EXP_FSST, 1, 0, 0, (uint8_t)(vec_c & 0xff), (uint8_t)(vec_c >> 8), (uint8_t)more, 0};
        put(hdr, 8);
        uint64_t b0 = 0;
        for (uint64_t k = v; k < stop2; ++k) b0 += comp[k].size();
        const uint32_t meta = 8 + 8 + 8 * (vec_c - 1);
        const uint32_t offs[2] = {meta, (uint32_t)(meta + b0)};
        put(offs, 8);
        uint32_t e0 = 0;
        for (uint64_t k = v; k < stop2; ++k) {
            if (k > v) {
                const uint32_t ep[2] = {e0, (uint32_t)(4096 * (k - v))};
                put(ep, 8);
            }
            e0 += (uint32_t)comp[k].size();
        }
        for (uint64_t k = v; k < stop2; ++k) put(comp[k].data(), comp[k].size());
        put(endsT.data() + v * 1024, 4096ull * vec_c);
        v = stop2;
        if (!more) break;
    }
    if (o.size() > cap) return -2;
    std::memcpy(out, o.data(), o.size());
    return (int64_t)o.size();
}

}  // namespace fls
