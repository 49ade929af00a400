// k_rowgroup.hip -- one persistent launch that decodes a whole row-group from a device-side work list.
//
// The host cuts every page of every column into tiles of up to 64 consecutive vectors (RgTile, rowgroup.hpp).
// Every wave takes tiles from one atomic counter until the list is exhausted and decodes a tile one vector at a
// time:
//   * the tile's per-vector metadata (bit width, FFOR base, entry point of the packed buffer, RLE run extents) is
//     loaded once, one vector per lane, and read per vector with v_readlane into scalar registers;
//   * the packed rows of vector k + 1 are loaded into registers (16-byte loads, all 64 lanes) while vector k is
//     decoded, then written to the wave's LDS slot;
//   * each thread produces a contiguous run of untransposed values (16 values; 2 x 8 for 8-bit) from the staged
//     rows with the vector's bit width as a run-time value (a 64/128-bit funnel shift per value), so one body
//     serves every bit width, and stores it with 16-byte stores.
// Every body is inlined into the kernel: the kernel makes no calls.
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

// 16-byte chunks of a vector's packed rows, loaded into registers ahead of their use.
template <unsigned N>
struct Prefetch {
    v4u c[N];
};
// Every lane issues exactly N loads (chunks past n16 re-read the vector's last chunk, or the page's first 16 bytes
// when the vector has none), so the number of memory operations in flight is the same on every path and the
// compiler waits for these loads without also waiting for the stores issued after them.
template <unsigned N>
__device__ __forceinline__ void prefetch_rows(Prefetch<N>& pf, const uint8_t* src, uint32_t n16, unsigned lane,
                                              const uint8_t* safe) {
    static_for<N>([&](auto k) {
        const uint32_t c = lane + 64u * decltype(k)::value;
        const uint8_t* a = n16 ? src + 16ull * (c < n16 ? c : n16 - 1) : safe;
        pf.c[k]          = *reinterpret_cast<const v4u_ua*>(a);
    });
}
// Writes all N prefetched chunks of every lane to LDS (chunks past n16 land in the wave's slot past the vector's
// rows, which nothing reads before they are overwritten): no branch separates the loads from these writes, so
// the compiler waits for each chunk's own load only, not for the stores issued after the loads. Rows past the N
// chunks per lane (64-bit vectors wider than 32 bits, src != nullptr) are copied from `src` now.
template <unsigned N>
__device__ __forceinline__ void stage_rows(const Prefetch<N>& pf, uint8_t* lds, uint32_t n16, unsigned lane,
                                           const uint8_t* src = nullptr) {
    static_for<N>([&](auto k) {
        *reinterpret_cast<v4u*>(lds + 16u * (lane + 64u * decltype(k)::value)) = pf.c[k];
    });
    if (src)
        for (uint32_t c = lane + 64u * N; c < n16; c += 64u)
            *reinterpret_cast<v4u*>(lds + 16u * c) = *reinterpret_cast<const v4u_ua*>(src + 16ull * c);
}

__device__ __forceinline__ uint32_t ep_at(const uint8_t* page, uint32_t arr_c, uint32_t v, uint32_t b) {
    return v ? reinterpret_cast<const u32_ua*>(page + 8 + 4 * arr_c)[(v - 1) * arr_c + b] : 0u;
}

// Accumulates values of T into 16-byte words and stores each word as soon as it is full, so a thread holds one
// word of output at a time. Values are put in order: put(k, x) for k = 0, 1, ... with k * sizeof(T) < bytes.
template <typename T>
struct WordOut {
    uint8_t* dst;
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
        if constexpr (e == per - 1) *reinterpret_cast<v4u*>(dst + 16 * (K / per)) = w;
    }
};

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

// --------------------------------------------------------------------------------------------------------------
// Integer pages: MODE 0 UNCOMPRESSED [values], 1 FFOR [bw][base][packed], 2 DELTA [lane bases][bw][base][packed]
// (include/fls/cor/exp/{uncompressed,ffor,delta}/*_exp.hpp), fixed-stride metadata.
// --------------------------------------------------------------------------------------------------------------
template <int W>
struct IntMap {
    static constexpr unsigned SZ = W / 8;
    // 16-byte chunks per lane prefetched ahead: all of a vector up to bw = 32 (128 * bw bytes); wider 64-bit
    // vectors load their remaining rows when they are staged
    static constexpr unsigned PF = W == 64 ? 4 : W / 8;
};

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
// Value of lane i at bit offset `bit`, fetched and extracted at once (for loops that use it right away).
template <int W>
__device__ __forceinline__ uint64_t lane_value(const uint8_t* lds, unsigned i, uint32_t bit, uint64_t mask) {
    return lane_extract<W>(lane_fetch<W>(lds, i, bit), bit, mask);
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
                                           unsigned t, uint8_t* dst) {
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
        uint8_t* const d = dst + (W / 8) * (W == 64 ? 64 * i + 16 * g : seg_start<32>(i) + 16 * g);
        WordOut<T>     o {d, {}};
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
        WordOut<T>  o {dst + 2 * seg_start<16>(t), {}};
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
        WordOut<T> o {dst + seg_start<8>(t), {}};
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

// Lane base(s) of the thread's lane(s) of vector v (DELTA pages: buffer 0 holds L lane bases of W bits).
template <int W>
__device__ __forceinline__ void load_lane_bases(const uint8_t* lbuf, unsigned t, uint64_t& lb0, uint64_t& lb1) {
    using T = typename uint_of<W>::type;
    if constexpr (W == 64) lb0 = load_t<T>(lbuf + 8 * (t % 16));
    else if constexpr (W == 32) lb0 = load_t<T>(lbuf + 4 * (t % 32));
    else if constexpr (W == 16) lb0 = load_t<T>(lbuf + 2 * t);
    else lb0 = lbuf[t], lb1 = lbuf[t + 64];
}

template <int W, int MODE>
__device__ __forceinline__ void rg_int(const uint8_t* page, const RgTile& tl, uint8_t* obase, uint8_t* lds,
                                       unsigned lane_in) {
    const unsigned lane = opaque_lane(lane_in);
    using T                = typename uint_of<W>::type;
    constexpr unsigned SZ  = W / 8;
    constexpr unsigned O   = MODE == 2 ? 1 : 0;                        // bit-width buffer
    constexpr unsigned P   = MODE == 0 ? 0 : (MODE == 1 ? 2 : 3);      // packed buffer
    constexpr unsigned NPF = IntMap<W>::PF;
    const uint32_t     nv  = tl.nv;
    // per-vector metadata, vector tl.v0 + lane in lane `lane`
    uint32_t m_bw = W, m_ep = 0;
    uint64_t m_base = 0;
    if (lane < nv) {
        const uint32_t v = tl.v0 + lane;
        if constexpr (MODE == 0) {
            m_ep = v * (128u * W);
        } else {
            const uint32_t b = page[tl.off[O] + v];
            m_bw             = b < (unsigned)W ? b : (unsigned)W;  // the host validated bw <= W
            m_base           = load_t<T>(page + tl.off[O + 1] + (uint64_t)v * SZ);
            m_ep             = ep_at(page, tl.arr_c, v, P);
        }
    }
    const uint8_t* pk = page + tl.off[P];
    Prefetch<NPF>  pf;
    uint64_t       lb0 = 0, lb1 = 0;
    {
        const uint32_t bw = __builtin_amdgcn_readfirstlane(m_bw), ep = __builtin_amdgcn_readfirstlane(m_ep);
        prefetch_rows(pf, pk + ep, 8 * bw, lane, page);
        if constexpr (MODE == 2) load_lane_bases<W>(page + tl.off[0] + 128ull * tl.v0, lane, lb0, lb1);
    }
    for (uint32_t k = 0; k < nv; ++k) {
        const uint32_t bw   = readlane_t(m_bw, k);
        const T        base = readlane_t((T)m_base, k);
        stage_rows(pf, lds, 8 * bw, lane, W == 64 && bw > 32 ? pk + readlane_t(m_ep, k) : nullptr);
        const uint64_t cb0 = lb0, cb1 = lb1;
        if (k + 1 < nv) {
            const uint32_t nbw = readlane_t(m_bw, k + 1), nep = readlane_t(m_ep, k + 1);
            prefetch_rows(pf, pk + nep, 8 * nbw, lane, page);
            if constexpr (MODE == 2) load_lane_bases<W>(page + tl.off[0] + 128ull * (tl.v0 + k + 1), lane, lb0, lb1);
        }
        wave_sync();
        uint8_t* dst = obase + (uint64_t)(tl.v0 + k) * (1024 * SZ);
        int_vector<W, MODE == 2>(lds, bw, base, cb0, cb1, lane, dst);
        wave_sync();
    }
}

// --------------------------------------------------------------------------------------------------------------
// RLE pages, DELTA_NO_PATCH positions (include/fls/cor/exp/rle/rle_exp.hpp:56-74): per vector [run values T]
// [pos lane bases u16 x 64][bw][base u16][packed u16 deltas]; out[i] = runs[pos[i]] (src/cor/lyt/vec.cpp:318-332).
// Positions past the vector's last run are clamped (err bit 1).
// --------------------------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void rg_rle(const uint8_t* page, const RgTile& tl, uint8_t* obase, uint8_t* lds,
                                       unsigned lane_in, uint32_t* err) {
    const unsigned lane = opaque_lane(lane_in);
    constexpr unsigned SZ = sizeof(T);
    const uint32_t     nv = tl.nv;
    uint32_t           m_bw = 0, m_ep = 0, m_r0 = 0, m_r1 = 0, m_base = 0;
    const uint32_t     run_bytes = tl.off[1] - tl.off[0];
    if (lane < nv) {
        const uint32_t v = tl.v0 + lane;
        const uint32_t b = page[tl.off[2] + v];
        m_bw             = b < 16u ? b : 16u;
        m_base           = *reinterpret_cast<const u16_ua*>(page + tl.off[3] + 2ull * v);
        m_ep             = ep_at(page, tl.arr_c, v, 4);
        m_r0             = ep_at(page, tl.arr_c, v, 0);
        m_r1             = v + 1 < tl.vec_c ? ep_at(page, tl.arr_c, v + 1, 0) : run_bytes;
    }
    uint8_t* const       runs_lds = lds + 128 * 17;
    const uint8_t* const pk       = page + tl.off[4];
    const uint8_t* const rv       = page + tl.off[0];
    Prefetch<2>          pf, pr;
    uint32_t             lb = 0;
    auto                 fetch = [&](uint32_t k) {
        const uint32_t bw = readlane_t(m_bw, k), ep = readlane_t(m_ep, k);
        const uint32_t r0 = readlane_t(m_r0, k), r1 = readlane_t(m_r1, k);
        prefetch_rows(pf, pk + ep, 8 * bw, lane, page);
        const uint32_t nb = r1 - r0;
        prefetch_rows(pr, rv + r0, nb <= kRgRunsLds ? (nb + 15) / 16 : 0u, lane, page);
        lb = *reinterpret_cast<const u16_ua*>(page + tl.off[1] + 128ull * (tl.v0 + k) + 2 * lane);
    };
    fetch(0);
    bool bad = false;
    for (uint32_t k = 0; k < nv; ++k) {
        const uint32_t bw = readlane_t(m_bw, k), base = readlane_t(m_base, k);
        const uint32_t r0 = readlane_t(m_r0, k), r1 = readlane_t(m_r1, k);
        const uint32_t nb = r1 - r0, n_runs = nb / SZ;
        const bool     in_lds = nb <= kRgRunsLds;
        stage_rows(pf, lds, 8 * bw, lane);
        stage_rows(pr, runs_lds, (nb + 15) / 16, lane);
        const uint32_t clb = lb;
        if (k + 1 < nv) fetch(k + 1);
        wave_sync();
        const uint32_t mask = (uint32_t)bw_mask<16>(bw), last = n_runs ? n_runs - 1 : 0;
        WordOut<T>     o {obase + (uint64_t)(tl.v0 + k) * (1024 * SZ) + SZ * seg_start<16>(lane), {}};
        RowPair<16>    pr[16];
        static_for<16>([&](auto r) { pr[r] = lane_fetch<16>(lds, lane, j_of_p<16>(decltype(r)::value) * bw); });
        batch_fence();
        uint32_t idx[16];
        uint32_t acc = clb;
        static_for<16>([&](auto r) {
            constexpr unsigned R = decltype(r)::value;
            acc                  = (acc + (uint32_t)lane_extract<16>(pr[R], j_of_p<16>(R) * bw, mask) + base) & 0xffffu;
            bad |= acc > last;
            idx[R] = acc > last ? last : acc;
        });
        static_for<2>([&](auto hh) {  // gathers in two batches of 8 values
            constexpr unsigned H = 8 * decltype(hh)::value;
            T                  x[8];
            if (in_lds) static_for<8>([&](auto r) { x[r] = reinterpret_cast<const T*>(runs_lds)[idx[H + r]]; });
            else static_for<8>([&](auto r) { x[r] = load_t<T>(rv + r0 + (uint64_t)idx[H + r] * SZ); });
            batch_fence();
            static_for<8>([&](auto r) { o.template put<H + decltype(r)::value>(x[r]); });
        });
        bad |= n_runs == 0;  // the host rejects such pages
        wave_sync();
    }
    if (bad) atomicOr(err, 1u);
}

// --------------------------------------------------------------------------------------------------------------
// Dictionary pages (include/fls/cor/exp/dict/dict_exp.hpp): u32 codes, VAR 1 [codes], VAR 2 [bw][base][packed],
// VAR 7 [lane bases][bw][base][packed]; out[i] = dict[code[i]] from the page's decoded dictionary (D-typed values,
// D = max(32, bits of T)). Codes past the dictionary are clamped (err bit 2).
// --------------------------------------------------------------------------------------------------------------
template <typename T, int VAR>
__device__ __forceinline__ void rg_dict(const uint8_t* page, const RgTile& tl, uint8_t* obase, uint8_t* lds,
                                        unsigned lane_in, uint32_t* err) {
    const unsigned lane = opaque_lane(lane_in);
    using D               = typename uint_of<(sizeof(T) < 4 ? 32 : 8 * sizeof(T))>::type;
    constexpr unsigned SZ = sizeof(T);
    constexpr unsigned O  = VAR == 7 ? 1 : 0;
    constexpr unsigned P  = VAR == 1 ? 0 : (VAR == 7 ? 3 : 2);
    const uint32_t     nv = tl.nv;
    const D*           dict = reinterpret_cast<const D*>(tl.dict);
    const uint32_t     last = tl.dict_n - 1;  // the host only builds dictionary tiles with dict_n > 0
    uint32_t           m_bw = 32, m_ep = 0, m_base = 0;
    if (lane < nv) {
        const uint32_t v = tl.v0 + lane;
        if constexpr (VAR == 1) {
            m_ep = v * 4096u;
        } else {
            const uint32_t b = page[tl.off[O] + v];
            m_bw             = b < 32u ? b : 32u;
            m_base           = *reinterpret_cast<const u32_ua*>(page + tl.off[O + 1] + 4ull * v);
            m_ep             = ep_at(page, tl.arr_c, v, P);
        }
    }
    const uint8_t* const pk = page + tl.off[P];
    Prefetch<4>          pf;
    uint64_t             lb = 0, unused = 0;
    {
        const uint32_t bw = __builtin_amdgcn_readfirstlane(m_bw), ep = __builtin_amdgcn_readfirstlane(m_ep);
        prefetch_rows(pf, pk + ep, 8 * bw, lane, page);
        if constexpr (VAR == 7) load_lane_bases<32>(page + tl.off[0] + 128ull * tl.v0, lane, lb, unused);
    }
    const unsigned i = lane % 32, h = lane / 32;
    bool           bad = false;
    for (uint32_t k = 0; k < nv; ++k) {
        const uint32_t bw = readlane_t(m_bw, k), base = readlane_t(m_base, k);
        stage_rows(pf, lds, 8 * bw, lane);
        const uint32_t clb = (uint32_t)lb;
        if (k + 1 < nv) {
            prefetch_rows(pf, pk + readlane_t(m_ep, k + 1), 8 * readlane_t(m_bw, k + 1), lane, page);
            if constexpr (VAR == 7) load_lane_bases<32>(page + tl.off[0] + 128ull * (tl.v0 + k + 1), lane, lb, unused);
        }
        wave_sync();
        const uint32_t mask = (uint32_t)bw_mask<32>(bw), hb = h * bw;
        auto           bit_of = [&](unsigned R) { return (4 * (R % 8) + 2 * (R / 8)) * bw + hb; };
        RowPair<32>    pr[16];
        static_for<16>([&](auto r) { pr[r] = lane_fetch<32>(lds, i, bit_of(decltype(r)::value)); });
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
            const uint32_t y   = __shfl_up(acc, 32, 64);
            const uint32_t add = clb + (h ? y : 0u);
            static_for<16>([&](auto r) { code[r] += add; });
        }
        T x[16];
        static_for<16>([&](auto r) {
            const uint32_t c = code[r];
            bad |= c > last;
            x[r] = (T)dict[c > last ? last : c];
        });
        batch_fence();
        WordOut<T> o {obase + (uint64_t)(tl.v0 + k) * (1024 * SZ) + SZ * (seg_start<32>(i) + 16 * h), {}};
        static_for<16>([&](auto r) { o.template put<decltype(r)::value>(x[r]); });
        wave_sync();
    }
    if (bad) atomicOr(err, 2u);
}

}  // namespace

__global__ __launch_bounds__(64 * kRgWaves) __attribute__((amdgpu_waves_per_eu(4))) void k_rowgroup(const uint8_t* __restrict__ enc, uint8_t* __restrict__ out,
                                                            const RgTile* __restrict__ tiles, uint32_t n_tiles,
                                                            uint32_t* __restrict__ counter, uint32_t* err,
                                                            uint32_t wave_lds) {
    extern __shared__ __attribute__((aligned(16))) uint8_t rg_lds[];
    const unsigned lane = threadIdx.x % 64;
    uint8_t* const lds  = rg_lds + (threadIdx.x / 64) * wave_lds;
    // Each wave claims its next tile while it decodes the current one, so the claim's round trip overlaps work.
    auto claim = [&]() {
        uint32_t id = 0;
        if (lane == 0) id = atomicAdd(counter, 1u);
        return __builtin_amdgcn_readfirstlane(id);  // lane 0's claim, in a scalar register: the loop stays uniform
    };
    uint32_t next = claim();
    for (;;) {
        const uint32_t id = next;
        if (id >= n_tiles) break;
        next = claim();
        const RgTile   tl   = tiles[id];
        const uint8_t* page = enc + tl.page;
        uint8_t*       ob   = out + tl.out;
        switch (tl.kind) {
        case RG_UNC8: rg_int<8, 0>(page, tl, ob, lds, lane); break;
        case RG_UNC16: rg_int<16, 0>(page, tl, ob, lds, lane); break;
        case RG_UNC32: rg_int<32, 0>(page, tl, ob, lds, lane); break;
        case RG_UNC64: rg_int<64, 0>(page, tl, ob, lds, lane); break;
        case RG_FFOR8: rg_int<8, 1>(page, tl, ob, lds, lane); break;
        case RG_FFOR16: rg_int<16, 1>(page, tl, ob, lds, lane); break;
        case RG_FFOR32: rg_int<32, 1>(page, tl, ob, lds, lane); break;
        case RG_FFOR64: rg_int<64, 1>(page, tl, ob, lds, lane); break;
        case RG_DELTA8: rg_int<8, 2>(page, tl, ob, lds, lane); break;
        case RG_DELTA16: rg_int<16, 2>(page, tl, ob, lds, lane); break;
        case RG_DELTA32: rg_int<32, 2>(page, tl, ob, lds, lane); break;
        case RG_DELTA64: rg_int<64, 2>(page, tl, ob, lds, lane); break;
        case RG_RLE8: rg_rle<uint8_t>(page, tl, ob, lds, lane, err); break;
        case RG_RLE16: rg_rle<uint16_t>(page, tl, ob, lds, lane, err); break;
        case RG_RLE32: rg_rle<uint32_t>(page, tl, ob, lds, lane, err); break;
        case RG_RLE64: rg_rle<uint64_t>(page, tl, ob, lds, lane, err); break;
        case RG_DICT1_8: rg_dict<uint8_t, 1>(page, tl, ob, lds, lane, err); break;
        case RG_DICT1_16: rg_dict<uint16_t, 1>(page, tl, ob, lds, lane, err); break;
        case RG_DICT1_32: rg_dict<uint32_t, 1>(page, tl, ob, lds, lane, err); break;
        case RG_DICT1_64: rg_dict<uint64_t, 1>(page, tl, ob, lds, lane, err); break;
        case RG_DICT2_8: rg_dict<uint8_t, 2>(page, tl, ob, lds, lane, err); break;
        case RG_DICT2_16: rg_dict<uint16_t, 2>(page, tl, ob, lds, lane, err); break;
        case RG_DICT2_32: rg_dict<uint32_t, 2>(page, tl, ob, lds, lane, err); break;
        case RG_DICT2_64: rg_dict<uint64_t, 2>(page, tl, ob, lds, lane, err); break;
        case RG_DICT7_8: rg_dict<uint8_t, 7>(page, tl, ob, lds, lane, err); break;
        case RG_DICT7_16: rg_dict<uint16_t, 7>(page, tl, ob, lds, lane, err); break;
        case RG_DICT7_32: rg_dict<uint32_t, 7>(page, tl, ob, lds, lane, err); break;
        case RG_DICT7_64: rg_dict<uint64_t, 7>(page, tl, ob, lds, lane, err); break;
        default:
            if (lane == 0) atomicOr(err, 8u);
            break;
        }
    }
}

__global__ void k_rowgroup_reset(uint32_t* counter) { *counter = 0u; }

unsigned rg_kind_bits(int kind) {
    if (kind < 0 || kind >= RG_KIND_COUNT) return 0;
    return 8u << (kind % 4);
}

unsigned rg_wave_lds(int kind) {
    if (kind < 0 || kind >= RG_KIND_COUNT) return 0;
    const unsigned W = rg_kind_bits(kind);
    if (kind <= RG_DELTA64) return 128 * W + 128;            // packed rows (bw <= W) + pad row
    if (kind <= RG_RLE64) return 128 * 17 + kRgRunsLds + 16;  // u16 positions (bw <= 16) + pad row, run values
    return 128 * 33;                                          // u32 codes (bw <= 32) + pad row
}

hipError_t launch_rowgroup(const uint8_t* enc, uint8_t* out, const RgTile* tiles, uint32_t n_tiles, uint32_t* counter,
                           uint32_t* err, unsigned wave_lds, unsigned blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_rowgroup_reset, dim3(1), dim3(1), 0, s, counter);
    hipLaunchKernelGGL(k_rowgroup, dim3(blocks), dim3(64 * kRgWaves), wave_lds * kRgWaves, s, enc, out, tiles, n_tiles,
                       counter, err, wave_lds);
    return hipGetLastError();
}

}  // namespace fls
