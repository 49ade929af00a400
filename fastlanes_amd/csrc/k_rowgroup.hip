// k_rowgroup.hip -- one persistent launch that decodes a whole row-group from a device-side work list.
//
// The host cuts every page of every column into tiles of up to 64 consecutive vectors whose input bytes (packed
// rows, lane bases, run values) are up to three contiguous ranges of the page, about 8 KiB in all (RgTile,
// rowgroup.hpp), and gives every wave of the launch a contiguous run of the list, balanced by bytes (a shared
// claim counter would serialize on its device-wide atomic: about 14 ns per claim). While a wave decodes a tile
// from its LDS slot, the next tile's ranges are already in flight into registers (8 x 16 bytes per lane) and
// its per-vector metadata (bit width, FFOR base, entry point, RLE run extent; one vector per lane) into a second
// set of registers, so each wave keeps a whole tile of input in flight. Values are produced from the staged rows
// with the vector's bit width as a run-time value (a 64/128-bit funnel shift per value), so one body serves every
// bit width, and stored with 16-byte stores. Every body is inlined into the kernel: the kernel makes no calls.
//
// Thread mappings (t = lane of the wave; W-bit values, L = 1024/W lanes of W values; value j of lane i sits at
// untransposed position seg_start<W>(i) + p with j = j_of_p<W>(p), fls_lanes.hpp):
//   W=64: lane i = t % 16, quarter q = t / 16, p in [16q, 16q+16)      (DELTA: prefix over the 4 quarters)
//   W=32: lane i = t % 32, half h = t / 32,    p in [16h, 16h+16)      (DELTA: prefix over the 2 halves)
//   W=16: lane i = t,                          p in [0, 16)
//   W=8 : lanes t and t + 64 (adjacent segments), p in [0, 8) each
// RLE positions are 16-bit (the W=16 mapping), dictionary codes 32-bit (the W=32 mapping).
// Reference semantics: unffor/unpack/rsum/untranspose (primitives/fls_generated/fallback/scalar_aav_uf1/*),
// ffor_decompress (src/cor/prm/ffor/ffor_decompress.cpp:13-32), rle_decompress (src/cor/prm/rle/rle_decompress.cpp:
// 14-38), dictionary gather (include/fls/encoder/assert_eq.hpp:257-328); the per-column loop it replaces is
// src/encoder/exp_col_decoder.cpp:13-26 behind src/reader/reader.cpp:19-50.
#include "../../include/fls_gpu.h"
#include "gpu_internal.hpp"
#include "page_kernels.hpp"
#include "rowgroup.hpp"

namespace fls {

namespace {

// Bits [s, s + 32) of the 64-bit value hi:lo (s < 32).
__device__ __forceinline__ uint32_t funnel32(uint32_t hi, uint32_t lo, uint32_t s) {
    return __builtin_amdgcn_alignbit(hi, lo, s);
}
// Bits [s, s + 64) of the 128-bit value hi:lo (s < 64).
__device__ __forceinline__ uint64_t funnel64(uint64_t hi, uint64_t lo, uint32_t s) {
    return s ? (lo >> s) | (hi << (64 - s)) : lo;
}

template <typename T>
__device__ __forceinline__ T readlane_t(T x, uint32_t k) {
    if constexpr (sizeof(T) == 8) {
        const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, k), hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), k);
        return (T)(((uint64_t)hi << 32) | lo);
    } else {
        return (T)__builtin_amdgcn_readlane((uint32_t)x, k);
    }
}

// 16-byte chunks of a tile's staged ranges, loaded into registers ahead of its decode.
struct Prefetch {
    v4u c[kRgPrefetch];
};

// Page and LDS-slot byte offsets of the tile's 16-byte chunk c: 16 c plus the delta of the range holding it. The
// deltas are added as differences under a condition (no branch, and no select between fields, which the compiler
// would turn into an indexed load from a stack copy of the tile).
__device__ __forceinline__ void chunk_at(const RgTile& t, uint32_t c, int32_t& src, int32_t& dst) {
    const uint32_t n0 = (t.rlen[0] + 15) / 16, n01 = n0 + (t.rlen[1] + 15) / 16;
    const bool     in1 = c >= n0, in2 = c >= n01;
    const int32_t  s0 = t.sdelta[0], s1 = t.sdelta[1] - s0, s2 = t.sdelta[2] - t.sdelta[1];
    const int32_t  l0 = t.ldelta[0], l1 = t.ldelta[1] - l0, l2 = t.ldelta[2] - t.ldelta[1];
    src = 16 * (int32_t)c + s0 + (in1 ? s1 : 0) + (in2 ? s2 : 0);
    dst = 16 * (int32_t)c + l0 + (in1 ? l1 : 0) + (in2 ? l2 : 0);
}

// Every lane issues exactly kRgPrefetch loads (lanes without a chunk re-read the page's first 16 bytes), so the
// number of memory operations in flight is the same on every path.
__device__ __forceinline__ void load_chunks(const uint8_t* page, const RgTile& t, unsigned lane, Prefetch& pf) {
    static_for<kRgPrefetch>([&](auto k) {
        const uint32_t c = lane + 64u * decltype(k)::value;
        int32_t        src, dst;
        chunk_at(t, c, src, dst);
        pf.c[k] = *reinterpret_cast<const v4u_ua*>(page + (c < t.chunks ? src : 0));
    });
}
// Writes the prefetched chunks to their places in the wave's LDS slot (lanes without a chunk write the slot's dump
// word: no branch separates the loads from these writes, so the compiler waits for the chunks' own loads only).
// Chunks past the registers' share (tiles larger than kRgStage bytes: single wide vectors) are copied now.
__device__ __forceinline__ void stage_chunks(const uint8_t* page, const RgTile& t, unsigned lane, const Prefetch& pf,
                                             uint8_t* lds, uint32_t wave_lds) {
    static_for<kRgPrefetch>([&](auto k) {
        const uint32_t c = lane + 64u * decltype(k)::value;
        int32_t        src, dst;
        chunk_at(t, c, src, dst);
        *reinterpret_cast<v4u*>(lds + (c < t.chunks ? (uint32_t)dst : wave_lds - kRgDump)) = pf.c[k];
    });
    if (t.chunks > 64u * kRgPrefetch)
        for (uint32_t c = lane + 64u * kRgPrefetch; c < t.chunks; c += 64u) {
            int32_t src, dst;
            chunk_at(t, c, src, dst);
            *reinterpret_cast<v4u*>(lds + dst) = *reinterpret_cast<const v4u_ua*>(page + src);
        }
}

// Tile `id` of the list, read through the constant address space: the address is uniform over the wave, so the
// 128 bytes arrive in scalar registers (s_load) and every field derived from them stays scalar.
__device__ __forceinline__ RgTile load_tile(const RgTile* tiles, uint32_t id) {
    const auto* p = reinterpret_cast<const __attribute__((address_space(4))) uint32_t*>(
        reinterpret_cast<uintptr_t>(tiles + id));
    union {
        RgTile   t;
        uint32_t w[sizeof(RgTile) / 4];
    } u;
    static_for<sizeof(RgTile) / 4>([&](auto k) { u.w[k] = p[k]; });
    return u.t;
}

__device__ __forceinline__ uint32_t ep_at(const uint8_t* page, uint32_t arr_c, uint32_t v, uint32_t b) {
    return v ? reinterpret_cast<const u32_ua*>(page + 8 + 4 * arr_c)[(v - 1) * arr_c + b] : 0u;
}

// Per-vector metadata of a tile, vector v0 + k in lane k (k < nv): bit width, entry point of range 0's rows
// (UNCOMPRESSED values / codes: their fixed offset), FFOR base, and for RLE the run values' byte extent.
struct TileMeta {
    uint32_t bw, ep, r0, r1;
    uint64_t base;
};
__device__ __forceinline__ uint64_t load_sz(const uint8_t* p, unsigned sz) {
    return sz == 8 ? (uint64_t)load_t<uint64_t>(p) : sz == 4 ? (uint64_t)load_t<uint32_t>(p)
                   : sz == 2 ? (uint64_t)load_t<uint16_t>(p) : (uint64_t)*p;
}
__device__ __forceinline__ void load_meta(const uint8_t* page, const RgTile& t, unsigned lane, TileMeta& m) {
    const unsigned fam = t.kind / 4, W = 8u << (t.kind % 4), SZ = W / 8;
    m = TileMeta {};
    if (lane >= t.nv) return;
    const uint32_t v = t.v0 + lane;
    switch (fam) {
    case 0: m.bw = W, m.ep = v * 128u * W; break;  // UNCOMPRESSED
    case 1:                                         // FFOR [bw][base][packed]
        m.bw = min((uint32_t)page[t.off[0] + v], W), m.base = load_sz(page + t.off[1] + (uint64_t)v * SZ, SZ);
        m.ep = ep_at(page, t.arr_c, v, 2);
        break;
    case 2:  // DELTA [lane bases][bw][base][packed]
        m.bw = min((uint32_t)page[t.off[1] + v], W), m.base = load_sz(page + t.off[2] + (uint64_t)v * SZ, SZ);
        m.ep = ep_at(page, t.arr_c, v, 3);
        break;
    case 3:  // RLE [runs][pos lane bases][bw][base u16][packed positions]
        m.bw   = min((uint32_t)page[t.off[2] + v], 16u);
        m.base = *reinterpret_cast<const u16_ua*>(page + t.off[3] + 2ull * v);
        m.ep   = ep_at(page, t.arr_c, v, 4);
        m.r0   = ep_at(page, t.arr_c, v, 0);
        m.r1   = v + 1 < t.vec_c ? ep_at(page, t.arr_c, v + 1, 0) : t.off[1] - t.off[0];
        break;
    case 4: m.bw = 32, m.ep = v * 4096u; break;  // dictionary, UNCOMPRESSED codes
    case 5:                                      // dictionary [bw][base][packed codes]
        m.bw = min((uint32_t)page[t.off[0] + v], 32u), m.base = *reinterpret_cast<const u32_ua*>(page + t.off[1] + 4ull * v);
        m.ep = ep_at(page, t.arr_c, v, 2);
        break;
    default:  // dictionary [lane bases][bw][base][packed codes]
        m.bw = min((uint32_t)page[t.off[1] + v], 32u), m.base = *reinterpret_cast<const u32_ua*>(page + t.off[2] + 4ull * v);
        m.ep = ep_at(page, t.arr_c, v, 3);
        break;
    }
}

// Accumulates values of T into 16-byte words and writes each word, as soon as it is full, to the wave's LDS output
// buffer (byte `off` of the vector, padded: lds_pad), so a thread holds one word of output at a time. Values are put
// in order: put(k, x) for k = 0, 1, ... with k * sizeof(T) < bytes.
template <typename T>
struct WordOut {
    uint8_t* ob;
    uint32_t off;
    v4u      w;
    template <unsigned K>
    __device__ __forceinline__ void put(T x) {
        constexpr unsigned per = 16 / sizeof(T), e = K % per;
        if constexpr (sizeof(T) == 8) {
            w[2 * e] = (uint32_t)x, w[2 * e + 1] = (uint32_t)((uint64_t)x >> 32);
        } else if constexpr (sizeof(T) == 4) {
            w[e] = (uint32_t)x;
        } else {
            constexpr unsigned d = e * sizeof(T) / 4, sh = 8 * (e * sizeof(T) % 4);
            w[d] = sh ? (w[d] | ((uint32_t)x << sh)) : (uint32_t)x;
        }
        if constexpr (e == per - 1) *reinterpret_cast<v4u*>(ob + lds_pad(off + 16 * (K / per))) = w;
    }
};

// Stores the decoded vector held in the wave's output buffer: lane-consecutive 16-byte words, so every store
// instruction writes 1 KiB of contiguous output (whole cache lines), bypassing the caches (the output is not
// read again by this launch).
template <unsigned VB>
__device__ __forceinline__ void vector_out(const uint8_t* ob, uint8_t* dst, unsigned lane) {
    wave_sync();
    static_for<VB / 1024>([&](auto k) {
        const uint32_t c = lane + 64u * decltype(k)::value;
        st16<true>(dst + 16 * c, *reinterpret_cast<const v4u*>(ob + lds_pad(16 * c)));
    });
    wave_sync();
}

// Ends a scheduling region: the compiler issues the staged-row reads of a batch of values together (one LDS round
// trip per batch instead of one per value) and does not move them into the neighbouring batch.
__device__ __forceinline__ void batch_fence() { __builtin_amdgcn_sched_barrier(0); }

// The wave's lane index as a value the compiler cannot treat as loop-invariant: every body derives its addresses
// from it inside its own case, instead of the compiler hoisting all bodies' lane arithmetic out of the tile loop
// (where it would stay live, in VGPRs, across every case of the kernel).
__device__ __forceinline__ unsigned opaque_lane(unsigned lane) {
    asm volatile("" : "+v"(lane));
    return lane;
}

// The two staged words a value needs: row w and row w + 1 of lane i (row w of lane i at lds + (w * L + i) * W/8,
// w = bit / W). Reading them is one LDS instruction for 32/64-bit rows; values are fetched in batches and extracted
// afterwards.
template <int W>
struct RowPair {
    using U = typename uint_of<W == 64 ? 64 : 32>::type;
    U lo, hi;
};
template <int W>
__device__ __forceinline__ RowPair<W> lane_fetch(const uint8_t* lds, unsigned i, uint32_t bit) {
    constexpr unsigned L = 1024 / W;
    const uint32_t     w = bit / W;
    if constexpr (W == 64) {
        const auto* r = reinterpret_cast<const uint64_t*>(lds);
        return {r[w * L + i], r[(w + 1) * L + i]};
    } else if constexpr (W == 32) {
        const auto* r = reinterpret_cast<const uint32_t*>(lds);
        return {r[w * L + i], r[(w + 1) * L + i]};
    } else if constexpr (W == 16) {
        const auto* r = reinterpret_cast<const uint16_t*>(lds);
        return {r[w * L + i], r[(w + 1) * L + i]};
    } else {
        return {lds[w * L + i], lds[(w + 1) * L + i]};
    }
}
// Bits [s, s + bw) of hi:lo, s = bit % W (run-time bw, mask = 2^bw - 1).
template <int W>
__device__ __forceinline__ uint64_t lane_extract(const RowPair<W>& p, uint32_t bit, uint64_t mask) {
    const uint32_t s = bit % W;
    if constexpr (W == 64) return funnel64(p.hi, p.lo, s) & mask;
    else if constexpr (W == 32) return funnel32(p.hi, p.lo, s) & (uint32_t)mask;
    else return (((p.hi << W) | p.lo) >> s) & (uint32_t)mask;
}
template <int W>
__device__ __forceinline__ uint64_t bw_mask(uint32_t bw) {
    return bw >= 64 ? ~0ull : ((1ull << bw) - 1);
}

// Decodes one staged vector and stores it at `dst` (the vector's first output byte). DELTA with a lane split over
// several threads (W = 64, 32) keeps the thread's 16 running sums, scans the threads' totals with shuffles and adds
// each thread's prefix before the stores.
template <int W, bool DELTA>
__device__ __forceinline__ void int_vector(const uint8_t* lds, uint32_t bw, uint64_t base, uint64_t lb0, uint64_t lb1,
                                           unsigned t, uint8_t* ob) {
    using T             = typename uint_of<W>::type;
    const uint64_t mask = bw_mask<W>(bw);
    if constexpr (W == 64 || W == 32) {
        constexpr unsigned LW = 64 / (W == 64 ? 4 : 2);  // lanes of the wave per lane group (16 or 32)
        constexpr unsigned B  = W == 64 ? 8 : 16;        // values per batch of staged-row reads
        const unsigned     i = t % LW, g = t / LW;       // lane, quarter (W = 64) or half (W = 32)
        const uint32_t     gb = (W == 64 ? 2 * (g & 1) + (g >> 1) : g) * bw;
        auto bit_of = [&](unsigned R) {
            const unsigned c = W == 64 ? 8 * (R % 8) + 4 * (R / 8) : 4 * (R % 8) + 2 * (R / 8);
            return c * bw + gb;
        };
        WordOut<T> o {ob, (W / 8) * (W == 64 ? 64 * i + 16 * g : seg_start<32>(i) + 16 * g), {}};
        T              x[DELTA ? 16 : 1];
        T              acc = 0;
        static_for<16 / B>([&](auto bb) {
            RowPair<W> pr[B];
            static_for<B>([&](auto r) { pr[r] = lane_fetch<W>(lds, i, bit_of(decltype(bb)::value * B + r)); });
            batch_fence();
            static_for<B>([&](auto r) {
                constexpr unsigned R = decltype(bb)::value * B + decltype(r)::value;
                const T            v = (T)(lane_extract<W>(pr[r], bit_of(R), mask) + base);
                if constexpr (DELTA) acc += v, x[R] = acc;
                else o.template put<R>(v);
            });
            batch_fence();
        });
        if constexpr (DELTA) {
            // exclusive prefix of the lane group's totals, plus the lane base
            T incl = acc;
            T y    = __shfl_up(incl, LW, 64);
            if (g >= 1) incl += y;
            if constexpr (W == 64) {
                y = __shfl_up(incl, 32, 64);
                if (g >= 2) incl += y;
            }
            const T add = (T)lb0 + (incl - acc);
            static_for<16>([&](auto r) { o.template put<decltype(r)::value>((T)(x[r] + add)); });
        }
    } else if constexpr (W == 16) {
        WordOut<T>  o {ob, 2 * seg_start<16>(t), {}};
        T           acc = (T)lb0;
        RowPair<16> pr[16];
        static_for<16>([&](auto r) { pr[r] = lane_fetch<16>(lds, t, j_of_p<16>(decltype(r)::value) * bw); });
        batch_fence();
        static_for<16>([&](auto r) {
            constexpr unsigned R = decltype(r)::value;
            T v = (T)(lane_extract<16>(pr[R], j_of_p<16>(R) * bw, mask) + base);
            if constexpr (DELTA) acc = (T)(acc + v), v = acc;
            o.template put<R>(v);
        });
    } else {
        WordOut<T> o {ob, seg_start<8>(t), {}};
        T          a0 = (T)lb0, a1 = (T)lb1;
        RowPair<8> pa[8], pb[8];
        static_for<8>([&](auto r) {
            pa[r] = lane_fetch<8>(lds, t, decltype(r)::value * bw);
            pb[r] = lane_fetch<8>(lds, t + 64, decltype(r)::value * bw);
        });
        batch_fence();
        static_for<8>([&](auto r) {
            constexpr unsigned R = decltype(r)::value;
            T v = (T)(lane_extract<8>(pa[R], R * bw, mask) + base);
            if constexpr (DELTA) a0 = (T)(a0 + v), v = a0;
            o.template put<R>(v);
        });
        static_for<8>([&](auto r) {
            constexpr unsigned R = decltype(r)::value;
            T v = (T)(lane_extract<8>(pb[R], R * bw, mask) + base);
            if constexpr (DELTA) a1 = (T)(a1 + v), v = a1;
            o.template put<8 + R>(v);
        });
    }
}

// Lane base(s) of the thread's lane(s) (DELTA pages: L lane bases of W bits per vector).
template <int W>
__device__ __forceinline__ void load_lane_bases(const uint8_t* lbuf, unsigned t, uint64_t& lb0, uint64_t& lb1) {
    using T = typename uint_of<W>::type;
    if constexpr (W == 64) lb0 = load_t<T>(lbuf + 8 * (t % 16));
    else if constexpr (W == 32) lb0 = load_t<T>(lbuf + 4 * (t % 32));
    else if constexpr (W == 16) lb0 = load_t<T>(lbuf + 2 * t);
    else lb0 = lbuf[t], lb1 = lbuf[t + 64];
}

// --------------------------------------------------------------------------------------------------------------
// Integer pages: MODE 0 UNCOMPRESSED [values], 1 FFOR [bw][base][packed], 2 DELTA [lane bases][bw][base][packed]
// (include/fls/cor/exp/{uncompressed,ffor,delta}/*_exp.hpp), fixed-stride metadata. Range 0 holds the tile's
// packed rows (vector k at ep_k - ep_0), range 1 the DELTA lane bases (128 bytes per vector).
// --------------------------------------------------------------------------------------------------------------
template <int W, int MODE>
__device__ __forceinline__ void rg_int(const RgTile& t, const TileMeta& m, uint8_t* obase, const uint8_t* lds,
                                       uint8_t* ob, unsigned lane_in) {
    using T               = typename uint_of<W>::type;
    constexpr unsigned SZ = W / 8;
    const unsigned     lane = opaque_lane(lane_in);
    const uint32_t     ep0  = readlane_t(m.ep, 0);
    for (uint32_t k = 0; k < t.nv; ++k) {
        const uint32_t bw   = readlane_t(m.bw, k);
        const T        base = readlane_t((T)m.base, k);
        uint64_t       lb0 = 0, lb1 = 0;
        if constexpr (MODE == 2) load_lane_bases<W>(lds + t.rlds[1] + 128 * k, lane, lb0, lb1);
        int_vector<W, MODE == 2>(lds + (readlane_t(m.ep, k) - ep0), bw, base, lb0, lb1, lane, ob);
        vector_out<1024 * SZ>(ob, obase + (uint64_t)(t.v0 + k) * (1024 * SZ), lane);
    }
}

// --------------------------------------------------------------------------------------------------------------
// RLE pages, DELTA_NO_PATCH positions (include/fls/cor/exp/rle/rle_exp.hpp:56-74): per vector [run values T]
// [pos lane bases u16 x 64][bw][base u16][packed u16 deltas]; out[i] = runs[pos[i]] (src/cor/lyt/vec.cpp:318-332).
// Range 0 holds the position rows, range 1 the run values, range 2 the position lane bases. Positions past the
// vector's last run are clamped (err bit 1).
// --------------------------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void rg_rle(const RgTile& t, const TileMeta& m, uint8_t* obase, const uint8_t* lds,
                                       uint8_t* ob, unsigned lane_in, uint32_t* err) {
    constexpr unsigned SZ   = sizeof(T);
    const unsigned     lane = opaque_lane(lane_in);
    const uint32_t     ep0 = readlane_t(m.ep, 0), r00 = readlane_t(m.r0, 0);
    bool               bad = false;
    for (uint32_t k = 0; k < t.nv; ++k) {
        const uint32_t bw = readlane_t(m.bw, k), base = readlane_t((uint32_t)m.base, k);
        const uint32_t r0 = readlane_t(m.r0, k), r1 = readlane_t(m.r1, k);
        const uint32_t n_runs = (r1 - r0) / SZ, last = n_runs ? n_runs - 1 : 0;
        const uint8_t* rows   = lds + (readlane_t(m.ep, k) - ep0);
        const T*       runs   = reinterpret_cast<const T*>(lds + t.rlds[1] + (r0 - r00));
        const uint32_t mask   = (uint32_t)bw_mask<16>(bw);
        WordOut<T>     o {ob, SZ * seg_start<16>(lane), {}};
        RowPair<16>    pr[16];
        static_for<16>([&](auto r) { pr[r] = lane_fetch<16>(rows, lane, j_of_p<16>(decltype(r)::value) * bw); });
        batch_fence();
        uint32_t idx[16];
        uint32_t acc = *reinterpret_cast<const uint16_t*>(lds + t.rlds[2] + 128 * k + 2 * lane);
        static_for<16>([&](auto r) {
            constexpr unsigned R = decltype(r)::value;
            acc                  = (acc + (uint32_t)lane_extract<16>(pr[R], j_of_p<16>(R) * bw, mask) + base) & 0xffffu;
            bad |= acc > last;
            idx[R] = acc > last ? last : acc;
        });
        static_for<2>([&](auto hh) {  // run values in two batches of 8
            constexpr unsigned H = 8 * decltype(hh)::value;
            T                  x[8];
            static_for<8>([&](auto r) { x[r] = runs[idx[H + r]]; });
            batch_fence();
            static_for<8>([&](auto r) { o.template put<H + decltype(r)::value>(x[r]); });
        });
        vector_out<1024 * SZ>(ob, obase + (uint64_t)(t.v0 + k) * (1024 * SZ), lane);
        bad |= n_runs == 0;  // the host rejects such pages
    }
    if (bad) atomicOr(err, 1u);
}

// --------------------------------------------------------------------------------------------------------------
// Dictionary pages (include/fls/cor/exp/dict/dict_exp.hpp): u32 codes, VAR 1 [codes], VAR 2 [bw][base][packed],
// VAR 7 [lane bases][bw][base][packed]; out[i] = dict[code[i]] from the page's decoded dictionary (D-typed values,
// D = max(32, bits of T)). Range 0 holds the code rows, range 1 the DELTA lane bases. Codes past the dictionary
// are clamped (err bit 2).
// --------------------------------------------------------------------------------------------------------------
template <typename T, int VAR>
__device__ __forceinline__ void rg_dict(const RgTile& t, const TileMeta& m, uint8_t* obase, const uint8_t* lds,
                                        uint8_t* ob, unsigned lane_in, uint32_t* err) {
    using D               = typename uint_of<(sizeof(T) < 4 ? 32 : 8 * sizeof(T))>::type;
    constexpr unsigned SZ = sizeof(T);
    const unsigned     lane = opaque_lane(lane_in);
    const auto*        dict = reinterpret_cast<const __attribute__((address_space(1))) D*>(t.dict);  // global
    const uint32_t     last = t.dict_n - 1;  // the host only builds dictionary tiles with dict_n > 0
    const uint32_t     ep0  = readlane_t(m.ep, 0);
    const unsigned     i = lane % 32, h = lane / 32;
    bool               bad = false;
    for (uint32_t k = 0; k < t.nv; ++k) {
        const uint32_t bw = readlane_t(m.bw, k), base = readlane_t((uint32_t)m.base, k);
        const uint8_t* rows = lds + (readlane_t(m.ep, k) - ep0);
        const uint32_t mask = (uint32_t)bw_mask<32>(bw), hb = h * bw;
        auto           bit_of = [&](unsigned R) { return (4 * (R % 8) + 2 * (R / 8)) * bw + hb; };
        RowPair<32>    pr[16];
        static_for<16>([&](auto r) { pr[r] = lane_fetch<32>(rows, i, bit_of(decltype(r)::value)); });
        batch_fence();
        uint32_t code[16];
        uint32_t acc = 0;
        static_for<16>([&](auto r) {
            constexpr unsigned R = decltype(r)::value;
            uint32_t c = (uint32_t)lane_extract<32>(pr[R], bit_of(R), mask) + base;
            if constexpr (VAR == 7) acc += c, c = acc;
            code[R] = c;
        });
        if constexpr (VAR == 7) {  // DELTA codes: the second half of a lane adds the first half's total
            const uint32_t lb  = *reinterpret_cast<const uint32_t*>(lds + t.rlds[1] + 128 * k + 4 * i);
            const uint32_t y   = __shfl_up(acc, 32, 64);
            const uint32_t add = lb + (h ? y : 0u);
            static_for<16>([&](auto r) { code[r] += add; });
        }
        T x[16];
        static_for<16>([&](auto r) {
            const uint32_t c = code[r];
            bad |= c > last;
            x[r] = (T)dict[c > last ? last : c];
        });
        batch_fence();
        WordOut<T> o {ob, SZ * (seg_start<32>(i) + 16 * h), {}};
        static_for<16>([&](auto r) { o.template put<decltype(r)::value>(x[r]); });
        vector_out<1024 * SZ>(ob, obase + (uint64_t)(t.v0 + k) * (1024 * SZ), lane);
    }
    if (bad) atomicOr(err, 2u);
}

}  // namespace

__global__ __launch_bounds__(64 * kRgWaves) void k_rowgroup(
    const uint8_t* __restrict__ enc, uint8_t* __restrict__ out, const RgTile* __restrict__ tiles,
    const uint32_t* __restrict__ wave_begin, uint32_t* __restrict__ err, uint32_t wave_lds, uint32_t out_lds) {
    extern __shared__ __attribute__((aligned(16))) uint8_t rg_lds[];
    // the wave index is uniform over the wave: readfirstlane tells the compiler so, and everything derived from it
    // (the run bounds, the tile structs, the switch over the tile kind) lives in scalar registers
    const unsigned lane = threadIdx.x % 64, wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    uint8_t* const lds  = rg_lds + wave * (wave_lds + out_lds);  // staged ranges
    uint8_t* const ob   = lds + wave_lds;                        // output buffer
    // this wave's run of the work list (the host balanced the runs by bytes; no tile is shared between waves)
    const uint32_t gw = blockIdx.x * (blockDim.x / 64) + wave;
    const uint32_t end = wave_begin[gw + 1];
    uint32_t       id  = wave_begin[gw];
    // software pipeline over the run: tile id + 1's struct, metadata and staged ranges are in flight while tile id
    // decodes from LDS
    RgTile   tn {};
    TileMeta mn {};
    Prefetch pf;
    if (id < end) {
        tn = load_tile(tiles, id);
        load_meta(enc + tn.page, tn, lane, mn);
        load_chunks(enc + tn.page, tn, lane, pf);
    }
    for (; id < end; ++id) {
        const RgTile   tc = tn;
        const TileMeta mc = mn;
        stage_chunks(enc + tc.page, tc, lane, pf, lds, wave_lds);
        if (id + 1 < end) {
            tn = load_tile(tiles, id + 1);
            load_meta(enc + tn.page, tn, lane, mn);
            load_chunks(enc + tn.page, tn, lane, pf);
        }
        wave_sync();
        uint8_t* obase = out + tc.out;
        switch (tc.kind) {
        case RG_UNC8: rg_int<8, 0>(tc, mc, obase, lds, ob, lane); break;
        case RG_UNC16: rg_int<16, 0>(tc, mc, obase, lds, ob, lane); break;
        case RG_UNC32: rg_int<32, 0>(tc, mc, obase, lds, ob, lane); break;
        case RG_UNC64: rg_int<64, 0>(tc, mc, obase, lds, ob, lane); break;
        case RG_FFOR8: rg_int<8, 1>(tc, mc, obase, lds, ob, lane); break;
        case RG_FFOR16: rg_int<16, 1>(tc, mc, obase, lds, ob, lane); break;
        case RG_FFOR32: rg_int<32, 1>(tc, mc, obase, lds, ob, lane); break;
        case RG_FFOR64: rg_int<64, 1>(tc, mc, obase, lds, ob, lane); break;
        case RG_DELTA8: rg_int<8, 2>(tc, mc, obase, lds, ob, lane); break;
        case RG_DELTA16: rg_int<16, 2>(tc, mc, obase, lds, ob, lane); break;
        case RG_DELTA32: rg_int<32, 2>(tc, mc, obase, lds, ob, lane); break;
        case RG_DELTA64: rg_int<64, 2>(tc, mc, obase, lds, ob, lane); break;
        case RG_RLE8: rg_rle<uint8_t>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_RLE16: rg_rle<uint16_t>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_RLE32: rg_rle<uint32_t>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_RLE64: rg_rle<uint64_t>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT1_8: rg_dict<uint8_t, 1>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT1_16: rg_dict<uint16_t, 1>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT1_32: rg_dict<uint32_t, 1>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT1_64: rg_dict<uint64_t, 1>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT2_8: rg_dict<uint8_t, 2>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT2_16: rg_dict<uint16_t, 2>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT2_32: rg_dict<uint32_t, 2>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT2_64: rg_dict<uint64_t, 2>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT7_8: rg_dict<uint8_t, 7>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT7_16: rg_dict<uint16_t, 7>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT7_32: rg_dict<uint32_t, 7>(tc, mc, obase, lds, ob, lane, err); break;
        case RG_DICT7_64: rg_dict<uint64_t, 7>(tc, mc, obase, lds, ob, lane, err); break;
        default:
            if (lane == 0) atomicOr(err, 8u);
            break;
        }
        wave_sync();
    }
}

unsigned rg_kind_bits(int kind) {
    if (kind < 0 || kind >= RG_KIND_COUNT) return 0;
    return 8u << (kind % 4);
}

unsigned rg_blocks_per_cu(unsigned slot_lds, unsigned waves) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_rowgroup, 64 * waves, slot_lds * waves) != hipSuccess)
        return 1;
    return n > 0 ? (unsigned)n : 1u;
}

hipError_t launch_rowgroup(const uint8_t* enc, uint8_t* out, const RgTile* tiles, const uint32_t* wave_begin,
                           uint32_t* err, unsigned wave_lds, unsigned out_lds, unsigned blocks, unsigned waves,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_rowgroup, dim3(blocks), dim3(64 * waves), (wave_lds + out_lds) * waves, s, enc, out, tiles,
                       wave_begin, err, wave_lds, out_lds);
    return hipGetLastError();
}

}  // namespace fls
