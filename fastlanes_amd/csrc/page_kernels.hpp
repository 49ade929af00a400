// page_kernels.hpp -- kernels that decode FastLanes pages (and raw FFOR streams) on gfx950.
//
// Page layout (src/cor/lyt/page/page.cpp:63-130, src/cor/lyt/page/entry_point.cpp:21-41): 8-byte header,
// arr_c u32 buffer offsets (from the page start), (vec_c - 1) * arr_c u32 entry points (vector 0's are 0),
// then the buffers. Vector v's slice of buffer b starts at page + off[b] + ep[v][b].
// Per-expression buffer lists follow the decompress RPNs in include/fls/cor/exp/*/*_exp.hpp.
//
// Launch shape: blocks of any multiple of 64 threads (up to 512); blocks stride over the vectors, so a grid smaller
// than the vector count is persistent. NT selects non-temporal stores.
#pragma once
#include "decode_core.hpp"

namespace fls {

constexpr unsigned kMaxBlock = 512;

__device__ __forceinline__ const uint8_t* vbuf(const uint8_t* page, unsigned arr_c, unsigned v, unsigned b) {
    const u32_ua* off = reinterpret_cast<const u32_ua*>(page + 8);
    uint32_t      ep  = v ? reinterpret_cast<const u32_ua*>(page + 8 + 4 * arr_c)[(v - 1) * arr_c + b] : 0u;
    return page + off[b] + ep;
}

// Calls f(integral_constant<BW>) for the runtime bw (bw > W calls nothing; the host validated every bw).
// Each case ends with an empty asm statement naming its bit width: the cases then never share a tail, so the
// compiler cannot sink their identical final stores into one block fed by per-case values (which keeps every
// case's registers live across the whole switch and spills).
template <int W, typename F>
__device__ __forceinline__ void dispatch_bw(unsigned bw, F&& f) {
    static_for<W + 1>([&](auto b) {
        if (bw == b) {
            f(b);
            asm volatile("; bw %0" ::"n"(decltype(b)::value));
        }
    });
}

// Grid-stride loop over vectors with TPV threads per vector; t is the thread's index within its vector.
#define FLS_VECTOR_LOOP(N, TPV)                                                                              \
    const unsigned t     = threadIdx.x % (TPV);                                                              \
    const uint64_t vstep = (uint64_t)gridDim.x * (blockDim.x / (TPV));                                       \
    for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x / (TPV)) + threadIdx.x / (TPV); v < (N); v += vstep)

template <int W>
constexpr unsigned lane_tpv() {
    return W == 64 ? 16 : 32;
}

// ---------------------------------------------------------------------------------------------------------------
// Raw FFOR / unpack stream: n_vec packed vectors back to back (128*BW bytes each), one base per vector or none.
// Output in the lane-major order of generated::unffor / unpack (16-byte column mapping).
// ---------------------------------------------------------------------------------------------------------------
template <int W, int BW, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_stream(const uint8_t* __restrict__ packed, const void* __restrict__ bases,
                                                      uint8_t* __restrict__ out, uint64_t n_vec) {
    using T = typename uint_of<W>::type;
    FLS_VECTOR_LOOP(n_vec, kTPV) {
        const uint64_t base = bases ? (uint64_t) reinterpret_cast<const T*>(bases)[v] : 0ull;
        ffor_lane_major<W, BW, NT>(packed + v * 128 * BW, base, t, out + v * 128 * W);
    }
}

// One vector of an integer page, untransposed: FFOR / UNCOMPRESSED (BW = W, base 0) / DELTA.
template <int W, int BW, bool DELTA, typename S>
__device__ __forceinline__ void ffor_untransposed(const uint8_t* pk, uint64_t base, uint64_t lane_base, unsigned t,
                                                  const S& dst) {
    if constexpr (W == 64) {
        decode64<BW, DELTA>(pk, base, lane_base, t, dst);
    } else {
        uint32_t V[W];
        lanes_p32<W, BW, DELTA>(pk, (uint32_t)base, (uint32_t)lane_base, t, V);
        store_segments32<W>(V, t, dst);
    }
}

// Per-wave LDS staging of decoded vectors. A wave decodes VPW = 64 / TPV vectors (one per TPV-thread group) into
// LDS, then all 64 lanes write them to HBM as consecutive 16-byte chunks: every store instruction writes whole
// 128-byte lines, which the lane-segment layout alone cannot do.
template <unsigned TPV, unsigned VB>
struct WaveStage {
    static constexpr unsigned VPW   = 64 / TPV;
    static constexpr unsigned PVB   = VB + VB / 16;  // padded bytes per vector
    static constexpr unsigned BYTES = VPW * PVB;     // LDS bytes per wave
};
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <unsigned TPV, unsigned VB, bool NT>
__device__ __forceinline__ void wave_copy_out(const uint8_t* wl, uint8_t* dst, uint32_t n_valid, unsigned lane) {
    using WS              = WaveStage<TPV, VB>;
    constexpr unsigned CV = VB / 16;  // 16-byte chunks per vector
    for (uint32_t c = lane; c < n_valid * CV; c += 64) {
        const uint32_t slot = c / CV, off = (c % CV) * 16;
        st16<NT>(dst + slot * VB + off, *reinterpret_cast<const v4u*>(wl + slot * WS::PVB + lds_pad(off)));
    }
}

// Predicate pushdown: instead of writing a decoded vector, compare its values with a constant and write the
// selection as the reference's SelectionDS bitmap (include/fls/expression/selection_ds.hpp: 16 uint64 words per
// 1024-value vector, bit i of word j = value 64j + i) plus the number of selected values (n_selected_values).
// Operators follow src/primitive/predicate/equal.cpp (==, >, <) plus >=, <=, !=; values compare in the column's
// own type (signed or unsigned).
enum SelOp : uint32_t { SEL_EQ = 0, SEL_GT = 1, SEL_LT = 2, SEL_GE = 3, SEL_LE = 4, SEL_NE = 5 };
// The host reduces every (operator, constant, type) to one test: map the value to an order-preserving unsigned key
// (xor `flip`: the sign bit for signed types; doubles: the IEEE total-order key with -0.0 folded onto 0.0), then
// `key - lo <= span` in the type's width, inverted by `inv` (!=); `none` fixes the result to `none_val` (> the
// maximum, < the minimum, a NaN constant), and a NaN value yields `nan_val` (true only for !=).
struct SelArgs {
    uint64_t* bitmap;  // 16 words per vector of the launch
    uint32_t* counts;  // one per vector of the launch
    uint64_t  lo, span, flip;
    uint32_t  inv, none, none_val, nan_val, dbl;
};

template <typename T, bool DBL>
__device__ __forceinline__ bool sel_pred(T x, const SelArgs& sa) {
    if constexpr (DBL && sizeof(T) == 8) {
        const uint64_t mag = x & 0x7fffffffffffffffull;
        if (mag > 0x7ff0000000000000ull) return sa.nan_val != 0;  // NaN
        const uint64_t z   = mag == 0 ? 0ull : x;                   // -0.0 -> 0.0
        const uint64_t key = (z >> 63) ? ~z : (z | 0x8000000000000000ull);
        const bool     in  = key - sa.lo <= sa.span;
        return sa.none ? sa.none_val != 0 : (in != (sa.inv != 0));
    } else {
        const T    key = (T)(x ^ (T)sa.flip);
        const bool in  = (T)(key - (T)sa.lo) <= (T)sa.span;
        return sa.none ? sa.none_val != 0 : (in != (sa.inv != 0));
    }
}

// Sink that turns decoded values straight into selection bits: every value the decode body stores is tested, its
// bit gathered in a register per 32-value bitmap word, and the word OR-ed into the vector's LDS bitmap when the
// body moves on to another word (a lane segment is contiguous, so that happens once per 32 values).
template <typename T, bool DBL>
struct PredSink {
    uint32_t*        bits;  // LDS: the vector's 32 bitmap words, cleared beforehand
    const SelArgs*   sa;
    mutable uint32_t acc = 0;
    mutable int32_t  cur = -1;
    __device__ __forceinline__ void add(uint32_t idx, T x) const {
        const int32_t w = (int32_t)(idx >> 5);
        if (w != cur) {
            if (cur >= 0 && acc) atomicOr(&bits[cur], acc);
            acc = 0, cur = w;
        }
        acc |= (uint32_t)sel_pred<T, DBL>(x, *sa) << (idx & 31);
    }
    __device__ __forceinline__ void flush() const {
        if (cur >= 0 && acc) atomicOr(&bits[cur], acc);
        acc = 0, cur = -1;
    }
    template <typename V>
    __device__ __forceinline__ void put_n(uint32_t off, V x) const {
        constexpr unsigned per = sizeof(V) / sizeof(T);
        union {
            V u;
            T e[per];
        } b;
        b.u = x;
        static_for<per>([&](auto k) { add(off / (uint32_t)sizeof(T) + decltype(k)::value, b.e[k]); });
    }
    __device__ __forceinline__ void put16(uint32_t off, v4u x) const { put_n(off, x); }
    __device__ __forceinline__ void put8(uint32_t off, v2u x) const { put_n(off, x); }
    template <typename U>
    __device__ __forceinline__ void put(uint32_t off, U x) const { put_n(off, x); }
};

// The wave's staged vectors (slots 0 .. n_valid-1, first vector v0) into bitmap words and counts: lane l reads
// values 64j + l (consecutive lanes, consecutive values), a ballot forms word j, lane j keeps it, and lanes 0-15
// store the vector's 16 words together.
template <typename T, unsigned TPV, unsigned VB, bool DBL>
__device__ __forceinline__ void wave_select_k(const uint8_t* wl, uint64_t v0, uint32_t n_valid, unsigned lane,
                                              const SelArgs& sa) {
    using WS = WaveStage<TPV, VB>;
    for (uint32_t slot = 0; slot < n_valid; ++slot) {
        const uint8_t* vb   = wl + slot * WS::PVB;
        uint64_t       mine = 0;
        uint32_t       cnt  = 0;
        static_for<16>([&](auto j) {
            const uint32_t idx = 64 * decltype(j)::value + lane;
            const T        x   = *reinterpret_cast<const T*>(vb + lds_pad(idx * (uint32_t)sizeof(T)));
            const uint64_t m   = __ballot(sel_pred<T, DBL>(x, sa));
            if (lane == decltype(j)::value) mine = m;
            cnt += (uint32_t)__popcll(m);
        });
        if (lane < 16) sa.bitmap[(v0 + slot) * 16 + lane] = mine;
        if (lane == 0) sa.counts[v0 + slot] = cnt;
    }
}
template <typename T, unsigned TPV, unsigned VB>
__device__ __forceinline__ void wave_select(const uint8_t* wl, uint64_t v0, uint32_t n_valid, unsigned lane,
                                            const SelArgs& sa) {
    if (sizeof(T) == 8 && sa.dbl) wave_select_k<T, TPV, VB, true>(wl, v0, n_valid, lane, sa);
    else wave_select_k<T, TPV, VB, false>(wl, v0, n_valid, lane, sa);
}

// End of a wave step: copy the staged vectors out, or (SEL) write their selection.
template <typename T, unsigned TPV, unsigned VB, bool NT, bool SEL>
__device__ __forceinline__ void wave_out(const uint8_t* wl, uint8_t* out, uint64_t v0, uint32_t n_valid, unsigned lane,
                                         const SelArgs& sa) {
    if constexpr (SEL) wave_select<T, TPV, VB>(wl, v0, n_valid, lane, sa);
    else wave_copy_out<TPV, VB, NT>(wl, out + v0 * VB, n_valid, lane);
}

// Loop over the kernel's vectors wave by wave; body(v, t, sink) decodes vector v (< N) into the sink.
template <unsigned TPV, unsigned VB, bool NT, typename F>
__device__ __forceinline__ void staged_vector_loop(uint64_t N, uint8_t* out, F&& body) {
    extern __shared__ __attribute__((aligned(16))) uint8_t fls_lds[];
    using WS              = WaveStage<TPV, VB>;
    const unsigned lane   = threadIdx.x % 64, wave = threadIdx.x / 64, waves = blockDim.x / 64;
    const unsigned t      = lane % TPV, slot = lane / TPV;
    uint8_t*       wl     = fls_lds + wave * WS::BYTES;
    const LdsSink  sink {wl + slot * WS::PVB};
    const uint64_t stride = (uint64_t)gridDim.x * waves * WS::VPW;
    for (uint64_t v0 = ((uint64_t)blockIdx.x * waves + wave) * WS::VPW; v0 < N; v0 += stride) {
        const uint64_t v = v0 + slot;
        if (v < N) body(v, t, sink);
        wave_sync();
        const uint64_t left = N - v0;
        wave_copy_out<TPV, VB, NT>(wl, out + v0 * VB, left < WS::VPW ? (uint32_t)left : WS::VPW, lane);
        wave_sync();
    }
}

template <int W>
__device__ __forceinline__ uint64_t lane_base_of(const uint8_t* p, unsigned t) {
    if constexpr (W == 64) return reinterpret_cast<const u64_ua*>(p)[t];
    else return reinterpret_cast<const u32_ua*>(p)[t];
}

// ---------------------------------------------------------------------------------------------------------------
// Integer pages. MODE 0: UNCOMPRESSED [values], 1: FFOR [bw][base][packed], 2: DELTA [lane bases][bw][base][packed]
// (include/fls/cor/exp/{uncompressed,ffor,delta}/*_exp.hpp). Output untransposed.
// ---------------------------------------------------------------------------------------------------------------
template <int W, int MODE, typename S>
__device__ __forceinline__ void int_page_vector(const uint8_t* page, uint32_t arr_c, uint64_t v, unsigned t, const S& dst) {
    using T = typename uint_of<W>::type;
    if constexpr (MODE == 0) {
        ffor_untransposed<W, W, false>(vbuf(page, arr_c, v, 0), 0, 0, t, dst);
    } else {
        constexpr unsigned o    = MODE == 2 ? 1 : 0;
        const unsigned     bw   = *vbuf(page, arr_c, v, o);
        const uint64_t     base = load_t<T>(vbuf(page, arr_c, v, o + 1));
        const uint64_t     lb   = MODE == 2 ? lane_base_of<W>(vbuf(page, arr_c, v, 0), t) : 0ull;
        const uint8_t*     pk   = vbuf(page, arr_c, v, o + 2);
        dispatch_bw<W>(bw, [&](auto BW) { ffor_untransposed<W, BW, MODE == 2>(pk, base, lb, t, dst); });
    }
}

// Page layout handed to the kernels by value: buffer offsets from the page start, buffer count, and whether the
// fixed-size metadata buffers sit at v * size (then their addresses need no entry-point load).
struct PageMeta {
    uint32_t off[6];
    uint32_t arr_c;
    uint32_t fixed;
};

// Per-vector metadata of an integer page: bit width, FFOR base, the thread's lane base (DELTA) and the entry point of
// the packed (or plain) value buffer. With fixed-stride metadata buffers the loads are independent of each other.
struct IntMeta {
    uint64_t base, lb;
    uint32_t bw, ep;
};

template <int W, int MODE>
__device__ __forceinline__ IntMeta int_meta(const uint8_t* page, const PageMeta& pm, uint64_t v, unsigned t) {
    using T              = typename uint_of<W>::type;
    constexpr unsigned o = MODE == 2 ? 1 : 0;            // bw buffer
    constexpr unsigned P = MODE == 0 ? 0 : o + 2;        // value buffer
    const u32_ua*      eps = reinterpret_cast<const u32_ua*>(page + 8 + 4 * pm.arr_c);
    IntMeta            m {};
    if constexpr (MODE == 0) {
        m.ep = pm.fixed ? (uint32_t)(v * (128 * W)) : (v ? eps[(v - 1) * pm.arr_c] : 0u);
        return m;
    } else {
        if (pm.fixed) {
            m.bw   = page[pm.off[o] + v];
            m.base = load_t<T>(page + pm.off[o + 1] + v * (W / 8));
            if constexpr (MODE == 2) m.lb = lane_base_of<W>(page + pm.off[0] + v * 128, t);
        } else {
            m.bw   = *vbuf(page, pm.arr_c, v, o);
            m.base = load_t<T>(vbuf(page, pm.arr_c, v, o + 1));
            if constexpr (MODE == 2) m.lb = lane_base_of<W>(vbuf(page, pm.arr_c, v, 0), t);
        }
        m.ep = v ? eps[(v - 1) * pm.arr_c + P] : 0u;
        return m;
    }
}

// Staged loop as staged_vector_loop, software-pipelined one step deep: the metadata of the vector a thread decodes
// next is loaded right after the current vector's values are decoded, so it arrives while the wave writes its
// output, and each step waits on one round of loads (the packed values) instead of three dependent ones.
// DIRECT (32/64-bit values): no LDS staging; each thread writes its lane's contiguous segment (128 or 512 bytes,
// whole lines) straight to HBM, so occupancy is bounded by registers instead of the staging buffers.
template <int W, int MODE, bool NT, bool DIRECT = false, bool SEL = false>
__global__ __launch_bounds__(kMaxBlock) void k_page_int(const uint8_t* __restrict__ page, PageMeta pm, uint32_t vec_c,
                                                        uint8_t* __restrict__ out, SelArgs sa) {
    extern __shared__ __attribute__((aligned(16))) uint8_t fls_lds[];
    constexpr unsigned TPV = lane_tpv<W>(), VB = 128 * W;
    constexpr unsigned P   = MODE == 0 ? 0 : (MODE == 2 ? 3 : 2);
    using WS               = WaveStage<TPV, VB>;
    const unsigned lane = threadIdx.x % 64, wave = threadIdx.x / 64, waves = blockDim.x / 64;
    const unsigned t = lane % TPV, slot = lane / TPV;
    uint8_t*       wl = fls_lds + wave * WS::BYTES;
    const LdsSink  sink {wl + slot * WS::PVB};
    const uint64_t N      = vec_c;
    const uint64_t stride = (uint64_t)gridDim.x * waves * WS::VPW;
    uint64_t       v0     = ((uint64_t)blockIdx.x * waves + wave) * WS::VPW;
    IntMeta        m {};
    if (v0 + slot < N) m = int_meta<W, MODE>(page, pm, v0 + slot, t);
    if constexpr (SEL && W <= 32) {
        // predicate pushdown: values become bits in registers; LDS holds only the wave's bitmaps (128 B per vector).
        // (64-bit values take the staged path below: their register form compiled to slower code.)
        using T         = typename uint_of<W>::type;
        uint32_t* wbits = reinterpret_cast<uint32_t*>(fls_lds) + wave * (WS::VPW * 32);
        for (; v0 < N; v0 += stride) {
            const uint64_t v = v0 + slot;
            for (unsigned k = t; k < 32; k += TPV) wbits[slot * 32 + k] = 0u;
            wave_sync();
            if (v < N) {
                const uint8_t* pk     = page + pm.off[P] + m.ep;
                auto           decode = [&](const auto& dst) {
                    if constexpr (MODE == 0) {
                        ffor_untransposed<W, W, false>(pk, 0, 0, t, dst);
                    } else {
                        dispatch_bw<W>(m.bw, [&](auto BW) { ffor_untransposed<W, BW, MODE == 2>(pk, m.base, m.lb, t, dst); });
                    }
                    dst.flush();
                };
                decode(PredSink<T, false> {wbits + slot * 32, &sa});
            }
            if (v + stride < N) m = int_meta<W, MODE>(page, pm, v + stride, t);
            wave_sync();
            // lane = 16 q + j writes word j of the wave's vector q, and the 16 lanes of a vector sum its count
            const unsigned q = lane / 16, j = lane % 16;
            const bool     ok = q < WS::VPW && v0 + q < N;
            const uint64_t w  = ok ? ((uint64_t)wbits[q * 32 + 2 * j + 1] << 32) | wbits[q * 32 + 2 * j] : 0ull;
            if (ok) sa.bitmap[(v0 + q) * 16 + j] = w;
            unsigned c = (unsigned)__popcll(w);
            for (unsigned d = 8; d >= 1; d >>= 1) c += __shfl_xor(c, d, 16);
            if (ok && j == 0) sa.counts[v0 + q] = c;
            wave_sync();
        }
        return;
    }
    for (; v0 < N; v0 += stride) {
        const uint64_t v = v0 + slot;
        if (v < N) {
            const uint8_t* pk     = page + pm.off[P] + m.ep;
            auto           decode = [&](const auto& dst) {
                if constexpr (MODE == 0) {
                    ffor_untransposed<W, W, false>(pk, 0, 0, t, dst);
                } else {
                    dispatch_bw<W>(m.bw, [&](auto BW) { ffor_untransposed<W, BW, MODE == 2>(pk, m.base, m.lb, t, dst); });
                }
            };
            if constexpr (DIRECT) decode(GlobalSink<NT> {out + v * VB});
            else decode(sink);
        }
        if (v + stride < N) m = int_meta<W, MODE>(page, pm, v + stride, t);
        if constexpr (!DIRECT) {
            wave_sync();
            const uint64_t left = N - v0;
            wave_out<typename uint_of<W>::type, TPV, VB, NT, SEL>(wl, out, v0, left < WS::VPW ? (uint32_t)left : WS::VPW,
                                                                  lane, sa);
            wave_sync();
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Patched FFOR pages, 32- and 64-bit (include/fls/cor/exp/ffor/ffor_exp.hpp:56-130): per vector
// [exceptions T][positions][bw][base][packed]. The vector is FFOR-decoded into the wave's LDS staging, then the
// exceptions overwrite their transposed positions k (held untransposed at orig_of(k)) before the copy-out.
// VAR 3 (S_PATCH): u16 count + u16 positions (src/cor/prm/patch/s_patch/s_patch_decompress.cpp:13-30).
// VAR 4 (B_PATCH): 16 u64 bitmap words, exceptions in bit order (src/cor/prm/patch/patch.cpp:71-83).
// VAR 2 (LL_PATCH): u16 count + u16 first position; a patched slot's FFOR value minus base plus one is the distance
// to the next exception (src/cor/prm/patch/ll_patch/ll_patch_decompress.cpp:13-42). The chain is serial: one thread
// walks it. A step that leaves the vector sets err bit 8 and ends the chain (the host validated S positions).
// ---------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned orig_of(unsigned k) {
    const unsigned b = (k >> 4) & 7;
    return (k >> 7) + 8 * (((b & 1) << 2) | (b & 2) | (b >> 2)) + 64 * (k & 15);
}

struct PatchMeta {
    uint64_t base, lb;                    // FFOR base; the thread's lane base (RSUM forms)
    uint32_t bw, ep_exc, ep_pos, ep_pk;
};

// Vector v's metadata of a patched FFOR-shaped page over CW-bit values: [lane bases (RSUM)][exceptions][positions]
// [bw][base][packed]; o = 1 with lane bases. Thread t of the vector's slot reads lane base t.
template <int CW, bool RSUM>
__device__ __forceinline__ PatchMeta patch_meta(const uint8_t* page, const PageMeta& pm, uint64_t v, unsigned t) {
    using T            = typename uint_of<CW>::type;
    constexpr unsigned o = RSUM ? 1 : 0;
    const u32_ua* eps  = reinterpret_cast<const u32_ua*>(page + 8 + 4 * pm.arr_c);
    const u32_ua* row  = eps + (v ? (v - 1) * pm.arr_c : 0);
    PatchMeta     m {};
    if (v) m.ep_exc = row[o], m.ep_pos = row[o + 1], m.ep_pk = row[o + 4];
    if (pm.fixed) {
        m.bw   = page[pm.off[o + 2] + v];
        m.base = load_t<T>(page + pm.off[o + 3] + v * (CW / 8));
        if constexpr (RSUM) m.lb = load_t<T>(page + pm.off[0] + 128 * v + t * (CW / 8));
    } else {
        const uint32_t e2 = v ? (uint32_t)row[o + 2] : 0u, e3 = v ? (uint32_t)row[o + 3] : 0u;
        m.bw   = page[pm.off[o + 2] + e2];
        m.base = load_t<T>(page + pm.off[o + 3] + e3);
        if constexpr (RSUM) m.lb = load_t<T>(page + pm.off[0] + (v ? (uint32_t)row[0] : 0u) + t * (CW / 8));
    }
    return m;
}

// RSUM over a vector staged untransposed in LDS (delta_exp.hpp:56-115 decompress: RSUM after the patch): lane t's
// values are the CW consecutive positions from seg_start<CW>(t); each becomes the lane base plus the running sum of
// the lane's differences up to it (gen/rsum_src.cpp). 16 bytes at a time.
template <int CW>
__device__ __forceinline__ void lds_rsum(uint8_t* vl, uint64_t lb, unsigned t) {
    using T              = typename uint_of<CW>::type;
    constexpr unsigned K = 16 / (CW / 8);  // values per 16-byte chunk
    const uint32_t     o = seg_start<CW>(t) * (CW / 8);
    T                  acc = (T)lb;
    static_for<CW / K>([&](auto q) {
        v4u* c = reinterpret_cast<v4u*>(vl + lds_pad(o + 16 * q));
        v4u  x = *c;
        if constexpr (CW == 64) {
            uint64_t a = (uint64_t)x[0] | ((uint64_t)x[1] << 32), b = (uint64_t)x[2] | ((uint64_t)x[3] << 32);
            acc += a, a = acc, acc += b, b = acc;
            x = v4u {(uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32)};
        } else {
            static_for<4>([&](auto j) {
                constexpr int k = decltype(j)::value;
                acc += x[k], x[k] = acc;
            });
        }
        *c = x;
    });
}

// Exceptions of one vector into its LDS staging `vl`, by the TPV threads of its slot (t = 0 .. TPV-1).
template <int W, int VAR, unsigned TPV>
__device__ __forceinline__ void apply_patches(const uint8_t* exc, const uint8_t* pos, uint64_t base, unsigned t,
                                              uint8_t* vl, uint32_t* err) {
    using T                = typename uint_of<W>::type;
    constexpr unsigned SZ  = W / 8;
    auto               at  = [&](unsigned k) { return reinterpret_cast<T*>(vl + lds_pad(orig_of(k) * SZ)); };
    if constexpr (VAR == 3) {
        const unsigned c = *reinterpret_cast<const u16_ua*>(pos);
        for (unsigned i = t; i < c; i += TPV) {
            const unsigned k = *reinterpret_cast<const u16_ua*>(pos + 2 + 2 * i);
            *at(k & 1023)    = load_t<T>(exc + i * SZ);
        }
    } else if constexpr (VAR == 4) {
        // thread t < 16 owns bitmap word t; its first exception index is the popcount of the words before it (an
        // exclusive scan over the slot's lanes; every lane of an active slot takes part)
        uint64_t       unit = t < 16 ? *reinterpret_cast<const u64_ua*>(pos + 8 * t) : 0ull;
        const unsigned own  = __builtin_popcountll(unit);
        unsigned       incl = own;
        static_for<4>([&](auto d) {
            const unsigned y = __shfl_up(incl, 1u << decltype(d)::value, TPV);
            if (t >= (1u << decltype(d)::value)) incl += y;
        });
        unsigned c = incl - own;
        while (unit) {
            const unsigned r = __builtin_ctzll(unit);
            *at(64 * t + r)  = load_t<T>(exc + c * SZ);
            ++c;
            unit &= unit - 1;
        }
    } else {
        if (t == 0) {
            // the exception values are loaded 16 at a time ahead of the walk: one memory wait per 16 links
            constexpr unsigned K   = 16;
            const unsigned     c   = *reinterpret_cast<const u16_ua*>(pos);
            T                  cur = *reinterpret_cast<const u16_ua*>(pos + 2);
            bool               ok  = true;
            for (unsigned i0 = 0; i0 < c && ok; i0 += K) {
                T e[K];
                static_for<K>([&](auto k) { e[k] = i0 + k < c ? load_t<T>(exc + (i0 + k) * SZ) : (T)0; });
                static_for<K>([&](auto k) {
                    if (ok && i0 + k < c) {
                        if (cur >= 1024) {
                            atomicOr(err, 8u);
                            ok = false;
                        } else {
                            T* slot          = at((unsigned)cur);
                            const T next_off = (T)(*slot - (T)base + 1);
                            *slot            = e[k];
                            cur              = (T)(cur + next_off);
                        }
                    }
                });
            }
        }
    }
}

// Dictionary codes of a staged vector (u32, untransposed, slot `cl`) into their values (slot `ol`): codes >= dict_n
// are clamped (err bit 2). TPV threads per vector, four codes at a time.
template <typename T, typename D, unsigned TPV>
__device__ __forceinline__ void lds_gather(const uint8_t* cl, uint8_t* ol, const D* __restrict__ dict, uint32_t dict_n,
                                           unsigned t, uint32_t* err) {
    const uint32_t last = dict_n - 1;
    bool           bad  = false;
    for (unsigned c4 = t; c4 < 256; c4 += TPV) {
        const v4u c = *reinterpret_cast<const v4u*>(cl + lds_pad(16 * c4));
        T         x[4];
        static_for<4>([&](auto j) {
            constexpr int k = decltype(j)::value;
            bad |= c[k] > last;
            x[k] = (T)dict[c[k] > last ? last : c[k]];
        });
        if constexpr (sizeof(T) == 8) {
            *reinterpret_cast<v4u*>(ol + lds_pad(32 * c4)) =
                v4u {(uint32_t)x[0], (uint32_t)(x[0] >> 32), (uint32_t)x[1], (uint32_t)(x[1] >> 32)};
            *reinterpret_cast<v4u*>(ol + lds_pad(32 * c4 + 16)) =
                v4u {(uint32_t)x[2], (uint32_t)(x[2] >> 32), (uint32_t)x[3], (uint32_t)(x[3] >> 32)};
        } else if constexpr (sizeof(T) == 4) {
            *reinterpret_cast<v4u*>(ol + lds_pad(16 * c4)) = v4u {x[0], x[1], x[2], x[3]};
        } else if constexpr (sizeof(T) == 2) {
            *reinterpret_cast<v2u*>(ol + lds_pad(8 * c4)) =
                v2u {(uint32_t)x[0] | ((uint32_t)x[1] << 16), (uint32_t)x[2] | ((uint32_t)x[3] << 16)};
        } else {
            *reinterpret_cast<uint32_t*>(ol + lds_pad(4 * c4)) =
                (uint32_t)x[0] | ((uint32_t)x[1] << 8) | ((uint32_t)x[2] << 16) | ((uint32_t)x[3] << 24);
        }
    }
    if (bad) atomicOr(err, 2u);
}

// Patched FFOR-shaped pages, pipelined like k_page_int: the CW-bit values of a vector are FFOR-decoded into the
// wave's LDS staging, the exceptions overwrite their positions, RSUM adds the lane bases (DELTA and dictionary
// DELTA_* forms), and dictionary pages (DICT, CW = 32) gather their T-typed values into a second staging area
// before the copy-out.
//   FFOR  [exceptions][positions][bw][base][packed]                 (include/fls/cor/exp/ffor/ffor_exp.hpp:56-130)
//   DELTA [lane bases][exceptions][positions][bw][base][packed]     (include/fls/cor/exp/delta/delta_exp.hpp:56-131)
//   DICT  either form over u32 codes                                (include/fls/cor/exp/dict/dict_exp.hpp:62-209)
// Exception values come from exc_base + the vector's entry point: the page's exception buffer, or the decoded HZL
// chain of the *_REC forms.
template <typename T, int CW, int VAR, bool RSUM, bool DICT, bool NT, bool SEL = false>
__global__ __launch_bounds__(kMaxBlock) void k_page_patch(const uint8_t* __restrict__ page, PageMeta pm,
                                                          uint32_t vec_c, uint8_t* __restrict__ out,
                                                          uint32_t* __restrict__ err, SelArgs sa,
                                                          const uint8_t* __restrict__ exc_base,
                                                          const typename uint_of<(8 * sizeof(T) < 32 ? 32 : 8 * sizeof(T))>::type* __restrict__ dict,
                                                          uint32_t dict_n) {
    extern __shared__ __attribute__((aligned(16))) uint8_t fls_lds[];
    static_assert(DICT || 8 * sizeof(T) == CW, "value pages decode T-typed values");
    constexpr unsigned TPV = lane_tpv<CW>(), VB = 128 * CW, OB = 1024 * sizeof(T);
    using WS               = WaveStage<TPV, VB>;
    using OS               = WaveStage<TPV, OB>;
    constexpr unsigned WB  = WS::BYTES + (DICT ? OS::BYTES : 0u);  // LDS bytes per wave
    const unsigned lane = threadIdx.x % 64, wave = threadIdx.x / 64, waves = blockDim.x / 64;
    const unsigned t = lane % TPV, slot = lane / TPV;
    uint8_t*       wl = fls_lds + wave * WB;
    uint8_t*       ow = DICT ? wl + WS::BYTES : wl;  // output staging
    uint8_t*       vl = wl + slot * WS::PVB;
    const LdsSink  sink {vl};
    const uint64_t N      = vec_c;
    const uint64_t stride = (uint64_t)gridDim.x * waves * WS::VPW;
    uint64_t       v0     = ((uint64_t)blockIdx.x * waves + wave) * WS::VPW;
    constexpr unsigned o  = RSUM ? 1 : 0;
    PatchMeta      m {};
    if (v0 + slot < N) m = patch_meta<CW, RSUM>(page, pm, v0 + slot, t);
    for (; v0 < N; v0 += stride) {
        const uint64_t  v   = v0 + slot;
        const PatchMeta cur = m;
        if (v < N) {
            const uint8_t* pk = page + pm.off[o + 4] + cur.ep_pk;
            dispatch_bw<CW>(cur.bw, [&](auto BW) { ffor_untransposed<CW, BW, false>(pk, cur.base, 0, t, sink); });
        }
        if (v + stride < N) m = patch_meta<CW, RSUM>(page, pm, v + stride, t);
        wave_sync();
        if (v < N)
            apply_patches<CW, VAR, TPV>(exc_base + cur.ep_exc, page + pm.off[o + 1] + cur.ep_pos, cur.base, t, vl,
                                        err);
        wave_sync();
        if constexpr (RSUM) {
            if (v < N) lds_rsum<CW>(vl, cur.lb, t);
            wave_sync();
        }
        if constexpr (DICT) {
            if (v < N && dict_n) lds_gather<T, typename uint_of<(8 * sizeof(T) < 32 ? 32 : 8 * sizeof(T))>::type, TPV>(
                vl, ow + slot * OS::PVB, dict, dict_n, t, err);
            wave_sync();
        }
        const uint64_t left = N - v0;
        wave_out<T, TPV, OB, NT, SEL>(ow, out, v0, left < WS::VPW ? (uint32_t)left : WS::VPW, lane, sa);
        wave_sync();
    }
}

// ---------------------------------------------------------------------------------------------------------------
// RLE pages (include/fls/cor/exp/rle/rle_exp.hpp:37-74). VAR 1: [run values][u16 positions, transposed];
// VAR 2: [run values][pos lane bases][bw][base u16][packed u16 deltas]. out[i] = runs[pos[i]]
// (src/cor/lyt/vec.cpp:318-332). Positions are clamped to the vector's run count; a clamp sets err bit 1.
// A thread owns one 32-bit column of the u16 position rows: two position lanes.
// ---------------------------------------------------------------------------------------------------------------
// Run value idx: from the page's (possibly unaligned) run buffer (R = uint8_t), or from an aligned copy (R = T).
template <typename T, typename R>
__device__ __forceinline__ T run_value(const R* runs, uint32_t idx) {
    if constexpr (sizeof(R) == 1) return load_t<T>(reinterpret_cast<const uint8_t*>(runs) + (uint64_t)idx * sizeof(T));
    else return runs[idx];
}

template <typename T, int BW, bool DELTA, typename S, typename R>
__device__ __forceinline__ void rle_vector(const uint8_t* packed, uint32_t base, uint32_t lane_base, unsigned t,
                                           const R* runs, uint32_t n_runs, const S& out, uint32_t* err) {
    uint32_t V[16];
    lanes_p32<16, BW, DELTA>(packed, base, lane_base, t, V);
    const uint32_t last = n_runs - 1;
    bool           bad  = false;
    static_for<2>([&](auto l) {
        T vals[16];
        static_for<16>([&](auto p) {
            uint32_t idx = (V[p] >> (16 * l)) & 0xffffu;
            bad |= idx > last;
            idx     = idx > last ? last : idx;
            vals[p] = run_value<T>(runs, idx);
        });
        store_run<T, 16, (8 * sizeof(T) < 16 ? 8 * sizeof(T) : 16)>(vals, sizeof(T) * seg_start<16>(2 * t + l), out);
    });
    if (bad) atomicOr(err, 1u);
}

// Run values of a vector with at most kRleLds runs are copied into LDS first and gathered from there.
constexpr unsigned kRleLds = 256;

template <typename T, int VAR, typename S, typename R>
__device__ __forceinline__ void rle_positions(const uint8_t* page, uint32_t arr_c, uint64_t v, unsigned t, const R* runs,
                                              uint32_t n_runs, const S& dst, uint32_t* err) {
    if constexpr (VAR == 1) {
        rle_vector<T, 16, false>(vbuf(page, arr_c, v, 1), 0, 0, t, runs, n_runs, dst, err);
    } else {
        const uint32_t lb   = reinterpret_cast<const u32_ua*>(vbuf(page, arr_c, v, 1))[t];
        const unsigned bw   = *vbuf(page, arr_c, v, 2);
        const uint32_t base = *reinterpret_cast<const u16_ua*>(vbuf(page, arr_c, v, 3));
        const uint8_t* pk   = vbuf(page, arr_c, v, 4);
        dispatch_bw<16>(bw, [&](auto BW) { rle_vector<T, BW, true>(pk, base, lb, t, runs, n_runs, dst, err); });
    }
}

template <typename T, int VAR, typename S>
__device__ __forceinline__ void rle_page_vector(const uint8_t* page, uint32_t arr_c, uint32_t vec_c, uint64_t v,
                                                unsigned t, const S& dst, uint32_t* err) {
    const uint8_t* runs = vbuf(page, arr_c, v, 0);
    // run count from the run buffer's extent: next vector's entry point, or the next buffer's offset
    const uint8_t* runs_end = v + 1 < vec_c ? vbuf(page, arr_c, v + 1, 0) : vbuf(page, arr_c, 0, 1);
    const uint32_t n_runs   = (uint32_t)((runs_end - runs) / sizeof(T));
    if (n_runs == 0) {  // the host rejects such pages; never read an empty run buffer
        atomicOr(err, 1u);
        return;
    }
    rle_positions<T, VAR>(page, arr_c, v, t, runs, n_runs, dst, err);
}

// Per-vector metadata of an RLE page: run-value bytes [r0, r1) of buffer 0, and for the positions: lane base
// (VAR 2), bit width, base, entry point of the position buffer.
struct RleMeta {
    uint32_t r0, r1, lb, bw, base, ep;
};

// Where a page's run values are: `base` holds them (T-typed), entry points of buffer 0 count 2^esh bytes per run and
// `bytes` is buffer 0's extent in those units. RLE pages: the page's own buffer 0 (esh = log2 sizeof(T)), or its
// decoded chain (the *_REC forms). DICT_RLE pages: run values looked up from the page's decoded run codes beforehand
// (esh = 2, u32 codes). `lbase`: the position lane bases' decoded chain (RLE DELTA_NO_PATCH_REC2), nullptr: the
// page's buffer 1.
struct RunSrc {
    const uint8_t* base;
    uint32_t       bytes, esh;
    const uint8_t* lbase = nullptr;
};

template <int VAR>
__device__ __forceinline__ RleMeta rle_meta(const uint8_t* page, const PageMeta& pm, uint64_t v, uint64_t N, unsigned t,
                                            uint32_t run_bytes, const uint8_t* lbase) {
    const u32_ua* eps = reinterpret_cast<const u32_ua*>(page + 8 + 4 * pm.arr_c);
    RleMeta       m {};
    m.r0 = v ? eps[(v - 1) * pm.arr_c] : 0u;
    m.r1 = v + 1 < N ? eps[v * pm.arr_c] : run_bytes;
    if constexpr (VAR == 1) {
        m.ep = pm.fixed ? (uint32_t)(v * 2048) : (v ? eps[(v - 1) * pm.arr_c + 1] : 0u);
    } else {
        const uint8_t* lb = lbase ? lbase : page + pm.off[1];
        if (pm.fixed) {
            m.lb   = reinterpret_cast<const u32_ua*>(lb + 128 * v)[t];
            m.bw   = page[pm.off[2] + v];
            m.base = *reinterpret_cast<const u16_ua*>(page + pm.off[3] + 2 * v);
        } else {
            m.lb   = reinterpret_cast<const u32_ua*>(lb + (v ? eps[(v - 1) * pm.arr_c + 1] : 0u))[t];
            m.bw   = *vbuf(page, pm.arr_c, v, 2);
            m.base = *reinterpret_cast<const u16_ua*>(vbuf(page, pm.arr_c, v, 3));
        }
        m.ep = v ? eps[(v - 1) * pm.arr_c + 4] : 0u;
    }
    return m;
}

// A wave decodes two vectors per step (32 threads each), pipelined like k_page_int. Per step: each vector's runs go
// to LDS when they fit, positions are decoded and the run values gathered into the wave's output staging, which is
// then written out.
// DIRECT (64-bit values): no output staging; a thread's 16-value run segments (128 bytes, whole lines) are stored
// straight to HBM, and LDS holds only the run values.
template <typename T, int VAR, bool NT, bool SEL = false, bool DIRECT = false>
__global__ __launch_bounds__(kMaxBlock) void k_page_rle(const uint8_t* __restrict__ page, PageMeta pm, uint32_t vec_c,
                                                        uint8_t* __restrict__ out, uint32_t* err, SelArgs sa,
                                                        RunSrc rs) {
    extern __shared__ __attribute__((aligned(16))) uint8_t fls_lds[];
    using WS              = WaveStage<32, 1024 * sizeof(T)>;
    constexpr unsigned RB = kRleLds * sizeof(T);
    constexpr unsigned P  = VAR == 1 ? 1 : 4;
    constexpr unsigned SB = DIRECT ? 0u : WS::BYTES;  // staging bytes per wave
    const unsigned     lane = threadIdx.x % 64, wave = threadIdx.x / 64, waves = blockDim.x / 64;
    const unsigned     t = lane % 32, slot = lane / 32;
    uint8_t*           wl    = fls_lds + wave * (SB + WS::VPW * RB);
    T*                 lruns = reinterpret_cast<T*>(wl + SB + slot * RB);
    const LdsSink      lsink {wl + slot * WS::PVB};
    const uint64_t     N      = vec_c;
    const uint64_t     stride = (uint64_t)gridDim.x * waves * WS::VPW;
    uint64_t           v0     = ((uint64_t)blockIdx.x * waves + wave) * WS::VPW;
    RleMeta            m {};
    if (v0 + slot < N) m = rle_meta<VAR>(page, pm, v0 + slot, N, t, rs.bytes, rs.lbase);
    for (; v0 < N; v0 += stride) {
        const uint64_t v      = v0 + slot;
        const uint8_t* runs   = rs.base + (uint64_t)(m.r0 >> rs.esh) * sizeof(T);
        const uint32_t n_runs = (m.r1 - m.r0) >> rs.esh;
        if (v < N && n_runs <= kRleLds)
            for (uint32_t i = t; i < n_runs; i += 32) lruns[i] = load_t<T>(runs + (uint64_t)i * sizeof(T));
        wave_sync();
        if (v < N) {
            const uint8_t* pk = page + pm.off[P] + m.ep;
            auto positions    = [&](auto runs_src) {
                auto body = [&](const auto& sink) {
                    if constexpr (VAR == 1) {
                        rle_vector<T, 16, false>(pk, 0, 0, t, runs_src, n_runs, sink, err);
                    } else {
                        dispatch_bw<16>(m.bw, [&](auto BW) {
                            rle_vector<T, BW, true>(pk, m.base, m.lb, t, runs_src, n_runs, sink, err);
                        });
                    }
                };
                if constexpr (DIRECT) body(GlobalSink<NT> {out + v * (1024 * sizeof(T))});
                else body(lsink);
            };
            if (n_runs == 0) atomicOr(err, 1u);  // the host rejects such pages; never read an empty run buffer
            else if (n_runs <= kRleLds) positions(static_cast<const T*>(lruns));
            else positions(runs);
        }
        if (v + stride < N) m = rle_meta<VAR>(page, pm, v + stride, N, t, rs.bytes, rs.lbase);
        wave_sync();
        if constexpr (!DIRECT) {
            const uint64_t left = N - v0;
            wave_out<T, 32, 1024 * sizeof(T), NT, SEL>(wl, out, v0, left < WS::VPW ? (uint32_t)left : WS::VPW, lane,
                                                       sa);
            wave_sync();
        }
    }
}

// Bits j_of_p<16>(0) .. j_of_p<16>(p) of a u16 position lane word: the differences up to value p in orig order.
__host__ __device__ constexpr uint32_t prefix_mask16(unsigned p) {
    uint32_t m = 0;
    for (unsigned q = 0; q <= p; ++q) m |= 1u << j_of_p<16>(q);
    return m;
}

// RLE pages with DELTA positions whose bit width is at most 1 -- every page the encoders write: a run index grows by
// 0 or 1 per value, and each lane's first difference is 0, so FFOR's base is 0 -- for 8/16-bit values. A wave
// decodes VPS vectors per step, one u16 position lane of each per thread: the lane's 16 differences are one packed u16 word, its positions are
// the lane base plus running bit counts (value p in orig order is bit j_of_p<16>(p)), and its 16 values -- a
// contiguous segment of the output -- are gathered from the vector's run values in LDS and stored straight to HBM
// (the wave's 64 segments tile the vector: whole lines). The last position bounds all 16 (positions grow along the
// lane): only a lane past the run count clamps (err bit 1).
template <typename T, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_page_rle_lanes(const uint8_t* __restrict__ page, PageMeta pm,
                                                              uint32_t vec_c, uint8_t* __restrict__ out, uint32_t* err,
                                                              RunSrc rs) {
    static_assert(sizeof(T) <= 2, "narrow values only");
    // vectors per wave step (two measured no faster for 8-bit values: the step's latency is not what limits them)
    constexpr int VPS = 1;
    extern __shared__ __attribute__((aligned(16))) uint8_t fls_lds[];
    // the wave index made provably uniform: the vector index and every per-vector address below stay in scalar
    // registers, and the per-vector metadata is read once per wave
    const unsigned lane = threadIdx.x % 64, wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64),
                   waves = blockDim.x / 64;
    T*             lruns = reinterpret_cast<T*>(fls_lds) + wave * 1024 * VPS;  // at most 1024 runs per vector
    const uint64_t N      = vec_c;
    const uint64_t stride = (uint64_t)gridDim.x * waves * VPS;
    uint64_t       v      = ((uint64_t)blockIdx.x * waves + wave) * VPS;  // the step's vectors: v .. v + VPS - 1
    // Pipelined two steps deep: while a step's VPS vectors are decoded, the run values of the next step's are in
    // flight into registers (the first 256 of each; the rare rest is loaded in its own step) and the metadata of the
    // step after.
    constexpr unsigned K = 4;  // run values per lane and vector held in registers
    auto meta = [&](uint64_t u) {
        RleMeta r = rle_meta<2>(page, pm, u, N, lane >> 1, rs.bytes, rs.lbase);
        r.lb      = (r.lb >> (16 * (lane & 1))) & 0xffffu;  // rle_meta reads the u32 of lanes 2t, 2t + 1
        r.r0      = __builtin_amdgcn_readfirstlane(r.r0);
        r.r1      = __builtin_amdgcn_readfirstlane(r.r1);
        r.bw      = __builtin_amdgcn_readfirstlane(r.bw);
        r.base    = __builtin_amdgcn_readfirstlane(r.base);
        r.ep      = __builtin_amdgcn_readfirstlane(r.ep);
        return r;
    };
    // the vector's packed word for this lane (bit width 0 or 1: row 0 holds lane l's u16 at 2 l)
    auto load_word = [&](const RleMeta& r) {
        return r.bw ? (uint32_t)*reinterpret_cast<const u16_ua*>(page + pm.off[4] + r.ep + 2 * lane) : 0u;
    };
    auto n_of    = [&](const RleMeta& r) { return (r.r1 - r.r0) >> rs.esh; };
    auto runs_of = [&](const RleMeta& r) { return rs.base + (uint64_t)(r.r0 >> rs.esh) * sizeof(T); };
    auto load_runs = [&](const RleMeta& r, T (&dst)[K]) {
        const uint8_t* rp = runs_of(r);
        const uint32_t n  = n_of(r);
        static_for<K>([&](auto k) {
            const uint32_t i = lane + 64 * decltype(k)::value;
            dst[k]           = i < n ? load_t<T>(rp + (uint64_t)i * sizeof(T)) : (T)0;
        });
    };
    RleMeta  m0[VPS] = {}, m1[VPS] = {};
    uint32_t w0[VPS] = {}, w1[VPS] = {};
    T        r0v[VPS][K] = {};
    static_for<VPS>([&](auto k) {
        if (v + k < N) m0[k] = meta(v + k), w0[k] = load_word(m0[k]), load_runs(m0[k], r0v[k]);
        if (v + stride + k < N) m1[k] = meta(v + stride + k), w1[k] = load_word(m1[k]);
    });
    for (; v < N; v += stride) {
        RleMeta  cur[VPS];
        uint32_t w[VPS];
        static_for<VPS>([&](auto k) {
            cur[k]             = m0[k];
            w[k]               = w0[k];
            const uint32_t n   = v + k < N ? n_of(cur[k]) : 0u;
            T*             lr  = lruns + 1024 * k;
            static_for<K>([&](auto q) {
                const uint32_t i = lane + 64 * decltype(q)::value;
                if (i < n) lr[i] = r0v[k][q];
            });
            if (n > 64 * K) {  // the runs beyond the registers (uniform: the vector's run count)
                const uint8_t* rp = runs_of(cur[k]);
                for (uint32_t i = lane + 64 * K; i < n && i < 1024; i += 64) lr[i] = load_t<T>(rp + (uint64_t)i * sizeof(T));
            }
        });
        // the next step's runs and words, the metadata of the step after
        static_for<VPS>([&](auto k) {
            m0[k] = m1[k], w0[k] = w1[k];
            if (v + stride + k < N) load_runs(m0[k], r0v[k]);
            if (v + 2 * stride + k < N) m1[k] = meta(v + 2 * stride + k), w1[k] = load_word(m1[k]);
        });
        wave_sync();
        static_for<VPS>([&](auto k) {
            if (v + k >= N) return;
            const uint32_t n_runs = n_of(cur[k]);
            const T*       lr     = lruns + 1024 * k;
            if (n_runs == 0) {  // the host rejects such pages; never read an empty run buffer
                if (lane == 0) atomicOr(err, 1u);
                return;
            }
            const uint32_t last  = (n_runs < 1024 ? n_runs : 1024) - 1;
            const uint32_t base  = cur[k].base & 0xffffu;
            // the largest position of the lane, without the u16 wrap (a lane that wraps clamps)
            const uint32_t top   = cur[k].lb + (uint32_t)__builtin_popcount(w[k]) + 16 * base;
            const bool     clamp = top > last;
            T              vals[16];
            if (base == 0 && !__any(clamp)) {
                // position p = lane base + the count of set bits among the lane word's first p + 1 differences:
                // one AND and one bit count (which adds the lane's first run address) per value
                const T* r0 = lr + cur[k].lb;
                static_for<16>([&](auto p) {
                    constexpr uint32_t M = prefix_mask16(decltype(p)::value);
                    vals[p] = r0[__builtin_popcount(w[k] & M)];
                });
            } else {
                uint32_t pos = cur[k].lb;
                static_for<16>([&](auto p) {
                    constexpr unsigned j = j_of_p<16>(p);
                    pos = (pos + ((w[k] >> j) & 1u) + base) & 0xffffu;
                    vals[p] = lr[pos < last ? pos : last];
                });
            }
            if (clamp) atomicOr(err, 1u);
            // a segment starts at a multiple of 16 values: 16-byte stores
            store_run<T, 16, 16>(vals, sizeof(T) * seg_start<16>(lane),
                                 GlobalSink<NT> {out + (v + k) * 1024 * sizeof(T)});
        });
        wave_sync();
    }
}

// DICT_RLE run values: rv[i] = dict[codes[i]] over a page's decoded run codes, so that the RLE kernel gathers
// values (include/fls/encoder/assert_eq.hpp:329-400 flattens the codes first and looks them up after; the lookup
// commutes with the gather). Codes >= dict_n are clamped (err bit 2).
template <typename T, typename D>
__global__ __launch_bounds__(256) void k_drle_values(const uint32_t* __restrict__ codes, uint64_t n,
                                                     const D* __restrict__ dict, uint32_t dict_n, T* __restrict__ rv,
                                                     uint32_t* err) {
    if (dict_n == 0) return;  // the host rejects such pages
    bool bad = false;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const uint32_t c = codes[i];
        bad |= c >= dict_n;
        rv[i] = (T)dict[c < dict_n ? c : dict_n - 1];
    }
    if (bad) atomicOr(err, 2u);
}

// ---------------------------------------------------------------------------------------------------------------
// Dictionary pages (include/fls/cor/exp/dict/dict_exp.hpp): u32 codes, VAR 1 [codes], VAR 2 [bw][base][packed],
// VAR 7 [lane bases][bw][base][packed]; out[i] = dict[code[i]]. Codes >= dict_n are clamped (err bit 2).
// The decoded dictionary holds D-typed values (D = max(32, bits of T)). One u32 code lane per thread.
// ---------------------------------------------------------------------------------------------------------------
template <typename T, typename D, int BW, bool DELTA, typename S>
__device__ __forceinline__ void dict_vector(const uint8_t* packed, uint32_t base, uint32_t lane_base, unsigned t,
                                            const D* __restrict__ dict, uint32_t dict_n, const S& out, uint32_t* err) {
    uint32_t V[32];
    lanes_p32<32, BW, DELTA>(packed, base, lane_base, t, V);
    const uint32_t last = dict_n - 1;
    bool           bad  = false;
    T              vals[32];
    static_for<32>([&](auto p) {
        uint32_t c = V[p];
        bad |= c > last;
        c       = c > last ? last : c;
        vals[p] = (T)dict[c];
    });
    store_run<T, 32, 16>(vals, sizeof(T) * seg_start<32>(t), out);
    if (bad) atomicOr(err, 2u);
}

template <typename T, typename D, int VAR, typename S>
__device__ __forceinline__ void dict_page_vector(const uint8_t* page, uint32_t arr_c, uint64_t v, unsigned t,
                                                 const D* __restrict__ dict, uint32_t dict_n, const S& dst,
                                                 uint32_t* err) {
    {
        if constexpr (VAR == 1) {
            dict_vector<T, D, 32, false>(vbuf(page, arr_c, v, 0), 0, 0, t, dict, dict_n, dst, err);
        } else {
            constexpr unsigned o    = VAR == 7 ? 1 : 0;
            const unsigned     bw   = *vbuf(page, arr_c, v, o);
            const uint32_t     base = *reinterpret_cast<const u32_ua*>(vbuf(page, arr_c, v, o + 1));
            const uint32_t     lb   = VAR == 7 ? reinterpret_cast<const u32_ua*>(vbuf(page, arr_c, v, 0))[t] : 0u;
            const uint8_t*     pk   = vbuf(page, arr_c, v, o + 2);
            dispatch_bw<32>(bw, [&](auto BW) {
                dict_vector<T, D, BW, VAR == 7>(pk, base, lb, t, dict, dict_n, dst, err);
            });
        }
    }
}

// Per-vector metadata of a dictionary page (u32 codes): bit width, FFOR base, lane base (VAR 7), entry point of
// the code buffer.
struct CodeMeta {
    uint32_t bw, base, lb, ep;
};

template <int VAR>
__device__ __forceinline__ CodeMeta code_meta(const uint8_t* page, const PageMeta& pm, uint64_t v, unsigned t) {
    const u32_ua* eps = reinterpret_cast<const u32_ua*>(page + 8 + 4 * pm.arr_c);
    CodeMeta      m {};
    if constexpr (VAR == 1) {
        m.ep = pm.fixed ? (uint32_t)(v * 4096) : (v ? eps[(v - 1) * pm.arr_c] : 0u);
    } else {
        constexpr unsigned o = VAR == 7 ? 1 : 0;
        if (pm.fixed) {
            m.bw   = page[pm.off[o] + v];
            m.base = *reinterpret_cast<const u32_ua*>(page + pm.off[o + 1] + 4 * v);
            if constexpr (VAR == 7) m.lb = reinterpret_cast<const u32_ua*>(page + pm.off[0] + 128 * v)[t];
        } else {
            m.bw   = *vbuf(page, pm.arr_c, v, o);
            m.base = *reinterpret_cast<const u32_ua*>(vbuf(page, pm.arr_c, v, o + 1));
            if constexpr (VAR == 7) m.lb = reinterpret_cast<const u32_ua*>(vbuf(page, pm.arr_c, v, 0))[t];
        }
        m.ep = v ? eps[(v - 1) * pm.arr_c + o + 2] : 0u;
    }
    return m;
}

struct HzlChunk {
    uint64_t src;    // byte offset of the packed data (FFOR) or raw bytes (tail) in the encoded buffer
    uint64_t dst;    // first value index in the decoded dictionary
    uint64_t base;   // FFOR base
    uint32_t bw;     // 255 = raw tail
    uint32_t exc_c;  // exceptions following the packed data
    uint32_t n;      // values in this chunk (1024, or the tail's count)
    uint32_t pad;
};

// Value k of an HZL chunk (src/cor/eng/hzl_de_cmp.cpp:28-77): a raw tail value, or FFOR value j of lane k % L
// (k = j * L + lane, the chunk's lane-major order) plus the chunk base.
template <typename D>
__device__ __forceinline__ D hzl_value(const uint8_t* in, const HzlChunk& ch, uint32_t k) {
    constexpr unsigned DW = 8 * sizeof(D), L = 1024 / DW;
    if (ch.bw == 255) return load_t<D>(in + (uint64_t)k * sizeof(D));
    const unsigned bw = ch.bw;
    D              v  = 0;
    if (bw) {
        const D        mask = bw >= DW ? (D)~(D)0 : (D)(((D)1 << bw) - 1);
        const unsigned j = k / L, lane = k % L, bit = j * bw, w = bit / DW, sh = bit % DW;
        v = load_t<D>(in + ((uint64_t)w * L + lane) * sizeof(D)) >> sh;
        if (sh + bw > DW) v |= load_t<D>(in + ((uint64_t)(w + 1) * L + lane) * sizeof(D)) << (DW - sh);
        v &= mask;
    }
    return (D)(v + (D)ch.base);
}

// The HZL payload of a page's dictionary, for kernels that decode it themselves.
struct HzlSrc {
    const uint8_t*  enc;     // the encoded buffer the chunk offsets refer to
    const HzlChunk* chunks;  // device table
    uint32_t        n_chunks;
    uint32_t        exc;     // some chunk carries exceptions (then the separate HZL kernels decode it)
};

// DICT_RLE run values in one pass over the page's run-code chunks: each code is unpacked from its HZL chunk and its
// value read straight from the packed dictionary chunk that holds it (code / 1024; the tail chunk follows the full
// ones). For chains without exceptions; codes >= dict_n are clamped (err bit 2).
template <typename T, typename D>
__global__ __launch_bounds__(256) void k_drle_fused(const uint8_t* __restrict__ enc, const HzlChunk* runs,
                                                    uint32_t n_runs, const HzlChunk* dict, uint32_t n_dict,
                                                    uint32_t dict_n, T* __restrict__ rv, uint32_t* err) {
    if (dict_n == 0 || n_dict == 0) return;  // the host rejects such pages
    bool bad = false;
    for (uint32_t c = blockIdx.x; c < n_runs; c += gridDim.x) {
        const HzlChunk ch  = runs[c];
        const uint8_t* in  = enc + ch.src;
        const uint32_t cnt = ch.bw == 255 ? ch.n : 1024u;
        for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) {
            uint32_t code = hzl_value<uint32_t>(in, ch, k);
            bad |= code >= dict_n;
            code              = code < dict_n ? code : dict_n - 1;
            const uint32_t di = (code >> 10) < n_dict ? (code >> 10) : n_dict - 1;
            const HzlChunk dc = dict[di];
            rv[ch.dst + k]    = (T)hzl_value<D>(enc + dc.src, dc, (uint32_t)(code - dc.dst));
        }
    }
    if (bad) atomicOr(err, 2u);
}

// Dictionary pages, pipelined like k_page_int. IN_LDS: the decoded dictionary is first copied into LDS (after the
// waves' staging areas) and the code gathers read LDS instead of L2; the host picks it when the dictionary fits.
// DIRECT (64-bit values): no output staging, segments stored straight to HBM (as k_page_rle); LDS holds only the
// dictionary.
// FROM_HZL (with IN_LDS): every block decodes the dictionary's HZL chunks straight into LDS, so no separate
// dictionary pass runs before the page (chunks without exceptions only).
template <typename T, typename D, int VAR, bool NT, bool IN_LDS, bool SEL = false, bool DIRECT = false,
          bool FROM_HZL = false>
__global__ __launch_bounds__(kMaxBlock) void k_page_dict(const uint8_t* __restrict__ page, PageMeta pm, uint32_t vec_c,
                                                         const D* __restrict__ dict, uint32_t dict_n,
                                                         uint8_t* __restrict__ out, uint32_t* err, SelArgs sa,
                                                         HzlSrc hz) {
    extern __shared__ __attribute__((aligned(16))) uint8_t fls_lds[];
    if (dict_n == 0) return;  // the host rejects such pages
    constexpr unsigned VB = 1024 * sizeof(T);
    constexpr unsigned P  = VAR == 1 ? 0 : (VAR == 7 ? 3 : 2);
    using WS              = WaveStage<32, VB>;
    constexpr unsigned SB = DIRECT ? 0u : WS::BYTES;  // staging bytes per wave
    const D* dp           = dict;
    if constexpr (IN_LDS) {
        D* ldict = reinterpret_cast<D*>(fls_lds + (blockDim.x / 64) * SB);
        if constexpr (FROM_HZL) {
            for (uint32_t c = 0; c < hz.n_chunks; ++c) {
                const HzlChunk ch  = hz.chunks[c];
                const uint8_t* in  = hz.enc + ch.src;
                const uint32_t cnt = ch.bw == 255 ? ch.n : 1024u;
                for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x)
                    if (ch.dst + k < dict_n) ldict[ch.dst + k] = hzl_value<D>(in, ch, k);
            }
        } else {
            for (uint32_t i = threadIdx.x; i < dict_n; i += blockDim.x) ldict[i] = dict[i];
        }
        __syncthreads();
        dp = ldict;
    }
    const unsigned lane = threadIdx.x % 64, wave = threadIdx.x / 64, waves = blockDim.x / 64;
    const unsigned t = lane % 32, slot = lane / 32;
    uint8_t*       wl = fls_lds + wave * SB;
    const LdsSink  lsink {wl + slot * WS::PVB};
    const uint64_t N      = vec_c;
    const uint64_t stride = (uint64_t)gridDim.x * waves * WS::VPW;
    uint64_t       v0     = ((uint64_t)blockIdx.x * waves + wave) * WS::VPW;
    CodeMeta       m {};
    if (v0 + slot < N) m = code_meta<VAR>(page, pm, v0 + slot, t);
    for (; v0 < N; v0 += stride) {
        const uint64_t v = v0 + slot;
        if (v < N) {
            const uint8_t* pk   = page + pm.off[P] + m.ep;
            auto           body = [&](const auto& sink) {
                if constexpr (VAR == 1) {
                    dict_vector<T, D, 32, false>(pk, 0, 0, t, dp, dict_n, sink, err);
                } else {
                    dispatch_bw<32>(m.bw, [&](auto BW) {
                        dict_vector<T, D, BW, VAR == 7>(pk, m.base, m.lb, t, dp, dict_n, sink, err);
                    });
                }
            };
            if constexpr (DIRECT) body(GlobalSink<NT> {out + v * VB});
            else body(lsink);
        }
        if (v + stride < N) m = code_meta<VAR>(page, pm, v + stride, t);
        if constexpr (!DIRECT) {
            wave_sync();
            const uint64_t left = N - v0;
            wave_out<T, 32, VB, NT, SEL>(wl, out, v0, left < WS::VPW ? (uint32_t)left : WS::VPW, lane, sa);
            wave_sync();
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------
// HZL dictionary payload chunks (src/cor/eng/hzl_de_cmp.cpp:28-77): an FFOR chunk decodes to 1024 values in array
// order (the lane-major order of unffor); the raw tail is copied; exceptions are patched by a second launch.
// ---------------------------------------------------------------------------------------------------------------

// One block per chunk, one thread per (value row j, lane): every thread extracts its value straight from the two
// packed words that hold it (word floor(j*bw/DW) and the next of its lane), so a chunk decodes with one round of
// loads instead of a serial walk over its rows. Output in array order: out[j * L + lane] = value + base.
template <int DW>
__global__ __launch_bounds__(1024) void k_hzl_chunks(const uint8_t* __restrict__ enc, const HzlChunk* chunks,
                                                     uint32_t n_chunks, uint8_t* __restrict__ dict) {
    using D                = typename uint_of<DW>::type;
    constexpr unsigned L   = 1024 / DW;
    for (uint32_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        const HzlChunk ch  = chunks[c];
        D*             out = reinterpret_cast<D*>(dict + ch.dst * sizeof(D));
        const uint8_t* in  = enc + ch.src;
        if (ch.bw == 255) {
            for (uint32_t k = threadIdx.x; k < ch.n; k += blockDim.x) out[k] = load_t<D>(in + (uint64_t)k * sizeof(D));
            continue;
        }
        const unsigned bw   = ch.bw;
        const D        base = (D)ch.base;
        const D        mask = bw >= DW ? (D)~(D)0 : (D)(((D)1 << bw) - 1);
        for (uint32_t k = threadIdx.x; k < 1024; k += blockDim.x) {
            const unsigned j = k / L, lane = k % L;
            D              v = 0;
            if (bw) {
                const unsigned bit = j * bw, w = bit / DW, sh = bit % DW;
                v                  = load_t<D>(in + ((uint64_t)w * L + lane) * sizeof(D)) >> sh;
                if (sh + bw > DW) v |= load_t<D>(in + ((uint64_t)(w + 1) * L + lane) * sizeof(D)) << (DW - sh);
                v &= mask;
            }
            out[k] = (D)(v + base);
        }
    }
}
template <int DW>
__global__ __launch_bounds__(256) void k_hzl_exceptions(const uint8_t* __restrict__ enc, const HzlChunk* chunks,
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
