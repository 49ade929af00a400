// page_kernels.hpp -- kernels that decode FastLanes pages (and raw FFOR streams) on gfx950.
//
// Page layout (src/cor/lyt/page/page.cpp:63-130, src/cor/lyt/page/entry_point.cpp:21-41): 8-byte header,
// arr_c u32 buffer offsets (from the page start), (vec_c - 1) * arr_c u32 entry points (vector 0's are 0),
// then the buffers. Vector v's slice of buffer b starts at page + off[b] + ep[v][b].
// Per-expression buffer lists follow the decompress RPNs in include/fls/cor/exp/*/*_exp.hpp.
#pragma once
#include "decode_core.hpp"

namespace fls {

constexpr unsigned kBlock = 256;

template <int W>
struct Geo {
    static constexpr unsigned TPV = W == 64 ? 16 : 32;  // threads per vector
    static constexpr unsigned VPB = kBlock / TPV;       // vectors per 256-thread block
};

__device__ __forceinline__ const uint8_t* vbuf(const uint8_t* page, unsigned arr_c, unsigned v, unsigned b) {
    const u32_ua* off = reinterpret_cast<const u32_ua*>(page + 8);
    uint32_t      ep  = v ? reinterpret_cast<const u32_ua*>(page + 8 + 4 * arr_c)[(v - 1) * arr_c + b] : 0u;
    return page + off[b] + ep;
}

// Calls f(integral_constant<BW>) for the runtime bw (uniform per vector; bw > W calls nothing).
template <int W, typename F>
__device__ __forceinline__ void dispatch_bw(unsigned bw, F&& f) {
    static_for<W + 1>([&](auto b) {
        if (bw == b) f(b);
    });
}

// FFOR / unpack (+ optional delta) of one vector, untransposed (LANE_MAJOR = false) or in the lane-major order
// of the reference's unffor (LANE_MAJOR = true).
template <int W, int BW, bool DELTA, bool LANE_MAJOR>
__device__ __forceinline__ void ffor_vector(const uint8_t* packed, uint64_t base, uint64_t lane_base, unsigned t,
                                            uint8_t* out) {
    if constexpr (W == 64) {
        decode64<BW, DELTA, LANE_MAJOR>(packed, base, lane_base, t, out);
    } else {
        uint32_t V[W];
        lanes_p32<W, BW, DELTA>(packed, (uint32_t)base, (uint32_t)lane_base, t, V);
        if constexpr (LANE_MAJOR) store_rows32_pmajor<W>(V, t, out);
        else store_segments32<W>(V, t, out);
    }
}

template <int W>
__device__ __forceinline__ uint64_t load_lane_base(const uint8_t* p, unsigned t) {
    if constexpr (W == 64) return reinterpret_cast<const u64_ua*>(p)[t];
    else return reinterpret_cast<const u32_ua*>(p)[t];
}

// ---------------------------------------------------------------------------------------------------------------
// Raw FFOR / unpack stream: n_vec packed vectors back to back (128*BW bytes each), one base per vector (or none).
// Output in the lane-major order of generated::unffor / unpack.
// ---------------------------------------------------------------------------------------------------------------
template <int W, int BW>
__global__ __launch_bounds__(kBlock) void k_unffor_stream(const uint8_t* __restrict__ packed,
                                                          const void* __restrict__ bases, uint8_t* __restrict__ out,
                                                          uint64_t n_vec) {
    using T          = typename uint_of<W>::type;
    const uint64_t v = (uint64_t)blockIdx.x * Geo<W>::VPB + threadIdx.x / Geo<W>::TPV;
    const unsigned t = threadIdx.x % Geo<W>::TPV;
    if (v >= n_vec) return;
    const uint64_t base = bases ? (uint64_t) reinterpret_cast<const T*>(bases)[v] : 0ull;
    ffor_vector<W, BW, false, true>(packed + v * 128 * BW, base, 0, t, out + v * 1024 * (W / 8));
}

// ---------------------------------------------------------------------------------------------------------------
// Integer pages. MODE 0: UNCOMPRESSED [values], 1: FFOR [bw][base][packed], 2: DELTA [lane bases][bw][base][packed]
// (include/fls/cor/exp/{uncompressed,ffor,delta}/*_exp.hpp). Output untransposed.
// ---------------------------------------------------------------------------------------------------------------
template <int W, int MODE>
__global__ __launch_bounds__(kBlock) void k_page_int(const uint8_t* __restrict__ page, uint32_t arr_c,
                                                     uint32_t vec_c, uint8_t* __restrict__ out) {
    using T          = typename uint_of<W>::type;
    const unsigned v = blockIdx.x * Geo<W>::VPB + threadIdx.x / Geo<W>::TPV;
    const unsigned t = threadIdx.x % Geo<W>::TPV;
    if (v >= vec_c) return;
    uint8_t* dst = out + (uint64_t)v * 1024 * (W / 8);
    if constexpr (MODE == 0) {
        ffor_vector<W, W, false, false>(vbuf(page, arr_c, v, 0), 0, 0, t, dst);
    } else {
        constexpr unsigned o         = MODE == 2 ? 1 : 0;
        const unsigned     bw        = *vbuf(page, arr_c, v, o);
        const uint64_t     base      = load_t<T>(vbuf(page, arr_c, v, o + 1));
        const uint64_t     lane_base = MODE == 2 ? load_lane_base<W>(vbuf(page, arr_c, v, 0), t) : 0ull;
        const uint8_t*     packed    = vbuf(page, arr_c, v, o + 2);
        dispatch_bw<W>(bw, [&](auto BW) { ffor_vector<W, BW, MODE == 2, false>(packed, base, lane_base, t, dst); });
    }
}

// ---------------------------------------------------------------------------------------------------------------
// RLE pages (include/fls/cor/exp/rle/rle_exp.hpp:37-74). VAR 1: [run values][u16 positions, transposed];
// VAR 2: [run values][pos lane bases][bw][base u16][packed u16 deltas]. out[i] = runs[pos[i]]
// (src/cor/lyt/vec.cpp:318-332). Positions are clamped to the vector's run count; a clamp sets err bit 1.
// ---------------------------------------------------------------------------------------------------------------
template <typename T, int BW, bool DELTA>
__device__ __forceinline__ void rle_vector(const uint8_t* packed, uint32_t base, uint32_t lane_base, unsigned t,
                                           const uint8_t* runs, uint32_t n_runs, T* out, uint32_t* err) {
    uint32_t V[16];
    lanes_p32<16, BW, DELTA>(packed, base, lane_base, t, V);
    const uint32_t last = n_runs ? n_runs - 1 : 0;
    bool           bad  = n_runs == 0;
    static_for<2>([&](auto l) {
        T vals[16];
        static_for<16>([&](auto p) {
            uint32_t idx = (V[p] >> (16 * l)) & 0xffffu;
            bad |= idx > last;
            idx     = idx > last ? last : idx;
            vals[p] = n_runs ? load_t<T>(runs + (uint64_t)idx * sizeof(T)) : T(0);
        });
        store_run<T, 16, (8 * sizeof(T) < 16 ? 8 * sizeof(T) : 16)>(vals, out + seg_start<16>(2 * t + l));
    });
    if (bad) atomicOr(err, 1u);
}

template <typename T, int VAR>
__global__ __launch_bounds__(kBlock) void k_page_rle(const uint8_t* __restrict__ page, uint32_t arr_c,
                                                     uint32_t vec_c, T* __restrict__ out, uint32_t* err) {
    const unsigned v = blockIdx.x * Geo<16>::VPB + threadIdx.x / Geo<16>::TPV;
    const unsigned t = threadIdx.x % Geo<16>::TPV;
    if (v >= vec_c) return;
    const uint8_t* runs = vbuf(page, arr_c, v, 0);
    // run count from the run buffer's extent: next vector's entry point, or the next buffer's offset
    const uint8_t* runs_end = v + 1 < vec_c ? vbuf(page, arr_c, v + 1, 0) : vbuf(page, arr_c, 0, 1);
    const uint32_t n_runs   = (uint32_t)((runs_end - runs) / sizeof(T));
    T*             dst      = out + (uint64_t)v * 1024;
    if constexpr (VAR == 1) {
        rle_vector<T, 16, false>(vbuf(page, arr_c, v, 1), 0, 0, t, runs, n_runs, dst, err);
    } else {
        const uint32_t lane_base = reinterpret_cast<const u32_ua*>(vbuf(page, arr_c, v, 1))[t];
        const unsigned bw        = *vbuf(page, arr_c, v, 2);
        const uint32_t base      = *reinterpret_cast<const u16_ua*>(vbuf(page, arr_c, v, 3));
        const uint8_t* packed    = vbuf(page, arr_c, v, 4);
        dispatch_bw<16>(bw, [&](auto BW) { rle_vector<T, BW, true>(packed, base, lane_base, t, runs, n_runs, dst, err); });
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Dictionary pages (include/fls/cor/exp/dict/dict_exp.hpp): u32 codes, VAR 1 [codes], VAR 2 [bw][base][packed],
// VAR 7 [lane bases][bw][base][packed]; out[i] = dict[code[i]]. Codes >= dict_n are clamped (err bit 2).
// The decoded dictionary holds DW-bit values (DW = max(32, bits of T)).
// ---------------------------------------------------------------------------------------------------------------
template <typename T, typename D, int BW, bool DELTA>
__device__ __forceinline__ void dict_vector(const uint8_t* packed, uint32_t base, uint32_t lane_base, unsigned t,
                                            const D* __restrict__ dict, uint32_t dict_n, T* out, uint32_t* err) {
    uint32_t V[32];
    lanes_p32<32, BW, DELTA>(packed, base, lane_base, t, V);
    const uint32_t last = dict_n ? dict_n - 1 : 0;
    bool           bad  = dict_n == 0;
    T              vals[32];
    static_for<32>([&](auto p) {
        uint32_t c = V[p];
        bad |= c > last;
        c       = c > last ? last : c;
        vals[p] = dict_n ? (T)dict[c] : T(0);
    });
    store_run<T, 32, 16>(vals, out + seg_start<32>(t));
    if (bad) atomicOr(err, 2u);
}

template <typename T, typename D, int VAR>
__global__ __launch_bounds__(kBlock) void k_page_dict(const uint8_t* __restrict__ page, uint32_t arr_c,
                                                      uint32_t vec_c, const D* __restrict__ dict, uint32_t dict_n,
                                                      T* __restrict__ out, uint32_t* err) {
    const unsigned v = blockIdx.x * Geo<32>::VPB + threadIdx.x / Geo<32>::TPV;
    const unsigned t = threadIdx.x % Geo<32>::TPV;
    if (v >= vec_c) return;
    T* dst = out + (uint64_t)v * 1024;
    if constexpr (VAR == 1) {
        dict_vector<T, D, 32, false>(vbuf(page, arr_c, v, 0), 0, 0, t, dict, dict_n, dst, err);
    } else {
        constexpr unsigned o         = VAR == 7 ? 1 : 0;
        const unsigned     bw        = *vbuf(page, arr_c, v, o);
        const uint32_t     base      = *reinterpret_cast<const u32_ua*>(vbuf(page, arr_c, v, o + 1));
        const uint32_t     lane_base = VAR == 7 ? reinterpret_cast<const u32_ua*>(vbuf(page, arr_c, v, 0))[t] : 0u;
        const uint8_t*     packed    = vbuf(page, arr_c, v, o + 2);
        dispatch_bw<32>(bw, [&](auto BW) {
            dict_vector<T, D, BW, VAR == 7>(packed, base, lane_base, t, dict, dict_n, dst, err);
        });
    }
}

// ---------------------------------------------------------------------------------------------------------------
// HZL dictionary payload chunks (src/cor/eng/hzl_de_cmp.cpp:28-77): FFOR chunk -> 1024 values in array order,
// then exceptions; the raw tail is copied by the host-side plan as a separate chunk kind.
// ---------------------------------------------------------------------------------------------------------------
struct HzlChunk {
    uint64_t src;    // byte offset of the packed data (FFOR) or raw bytes (tail) in the encoded buffer
    uint64_t dst;    // first value index in the decoded dictionary
    uint64_t base;   // FFOR base
    uint32_t bw;     // 255 = raw tail
    uint32_t exc_c;  // exceptions following the packed data
    uint32_t n;      // values in this chunk (1024, or the tail's count)
    uint32_t pad;
};

template <int DW>
__global__ __launch_bounds__(kBlock) void k_hzl_chunks(const uint8_t* __restrict__ enc, const HzlChunk* chunks,
                                                       uint32_t n_chunks, uint8_t* __restrict__ dict) {
    using D = typename uint_of<DW>::type;
    // one vector per Geo<DW>::TPV threads; chunks are processed by a block-stride loop
    const unsigned sub = threadIdx.x / Geo<DW>::TPV, t = threadIdx.x % Geo<DW>::TPV;
    for (uint32_t c = blockIdx.x * Geo<DW>::VPB + sub; c < n_chunks; c += gridDim.x * Geo<DW>::VPB) {
        const HzlChunk ch  = chunks[c];
        D*             out = reinterpret_cast<D*>(dict) + ch.dst;
        if (ch.bw == 255) {
            for (uint32_t k = t; k < ch.n; k += Geo<DW>::TPV) out[k] = load_t<D>(enc + ch.src + (uint64_t)k * sizeof(D));
            continue;
        }
        dispatch_bw<DW>(ch.bw, [&](auto BW) {
            ffor_vector<DW, BW, false, true>(enc + ch.src, ch.base, 0, t, reinterpret_cast<uint8_t*>(out));
        });
    }
}
// Exceptions are applied after all chunks are decoded (a second launch), so no intra-chunk ordering is needed.
template <int DW>
__global__ __launch_bounds__(kBlock) void k_hzl_exceptions(const uint8_t* __restrict__ enc, const HzlChunk* chunks,
                                                           uint32_t n_chunks, uint8_t* __restrict__ dict) {
    using D = typename uint_of<DW>::type;
    for (uint32_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        const HzlChunk ch = chunks[c];
        if (ch.bw == 255 || ch.exc_c == 0) continue;
        const uint8_t* pos = enc + ch.src + 128ull * ch.bw;
        const uint8_t* val = pos + 2ull * ch.exc_c;
        D*             out = reinterpret_cast<D*>(dict) + ch.dst;
        for (uint32_t e = threadIdx.x; e < ch.exc_c; e += blockDim.x) {
            const uint32_t p = *reinterpret_cast<const u16_ua*>(pos + 2ull * e);
            if (p < 1024) out[p] = load_t<D>(val + (uint64_t)e * sizeof(D));
        }
    }
}

}  // namespace fls
