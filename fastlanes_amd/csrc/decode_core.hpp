// decode_core.hpp -- per-vector decode bodies (FFOR / unpack, delta, RLE gather, dictionary gather) for gfx950.
//
// A vector is decoded by a group of TPV threads (32 for W <= 32, 16 for W = 64); `t` is the thread's index in
// that group. Every body is fully unrolled for a compile-time bit width BW, keeps the packed rows in VGPRs and
// writes the untransposed output segment of each of its lanes (see fls_lanes.hpp for the layout algebra).
// Reference semantics followed:
//   unffor   primitives/fls_generated/fallback/scalar_aav_uf1/fallback_scalar_aav_1024_uf1_unffor_src.cpp
//   rsum     .../fallback_scalar_aav_1024_uf1_rsum_src.cpp (lane base + running sum, wrap mod 2^W)
//   untranspose_i  .../fallback_scalar_aav_1024_uf1_untranspose_src.cpp, src/encoder/materializer.cpp:13-21
//   RLE gather     src/cor/lyt/vec.cpp:318-332;  dictionary gather  include/fls/encoder/assert_eq.hpp:257-328
#pragma once
#include <type_traits>
#include <utility>

#include "fls_lanes.hpp"

namespace fls {

template <int W>
struct uint_of;
template <>
struct uint_of<8> {
    using type = uint8_t;
};
template <>
struct uint_of<16> {
    using type = uint16_t;
};
template <>
struct uint_of<32> {
    using type = uint32_t;
};
template <>
struct uint_of<64> {
    using type = uint64_t;
};

// ---------------------------------------------------------------------------------------------------------------
// Row loads. Page buffers carry no alignment guarantee (a page is a byte stream), so rows are read through
// alignment-1 types; gfx950 executes these as ordinary dword loads.
// ---------------------------------------------------------------------------------------------------------------
template <int BW>
__device__ __forceinline__ void load_rows32(const uint8_t* packed, unsigned t, uint32_t (&R)[BW > 0 ? BW : 1]) {
    const u32_ua* p = reinterpret_cast<const u32_ua*>(packed);
    static_for<BW>([&](auto w) { R[w] = __builtin_nontemporal_load(&p[w * 32 + t]); });
}
template <int BW>
__device__ __forceinline__ void load_rows64(const uint8_t* packed, unsigned t, uint64_t (&R)[BW > 0 ? BW : 1]) {
    const u64_ua* p = reinterpret_cast<const u64_ua*>(packed);
    static_for<BW>([&](auto w) { R[w] = __builtin_nontemporal_load(&p[w * 16 + t]); });
}

// ---------------------------------------------------------------------------------------------------------------
// Lane values in untransposed (p) order. V[p] for W <= 32 holds the G = 32/W lanes of the thread's column.
// base: FFOR base (added mod 2^W to every unpacked value); lane_base: the column's lanes' delta bases.
// ---------------------------------------------------------------------------------------------------------------
template <int W, int BW, bool DELTA>
__device__ __forceinline__ void lanes_p32(const uint8_t* packed, uint32_t base, uint32_t lane_base, unsigned t,
                                          uint32_t (&V)[W]) {
    uint32_t R[BW > 0 ? BW : 1];
    load_rows32<BW>(packed, t, R);
    const uint32_t brep = rep<W>(base);
    uint32_t       acc  = lane_base;
    static_for<W>([&](auto p) {
        constexpr unsigned j = j_of_p<W>(p);
        uint32_t           v = swar_add<W>(extract32<W, BW, j>(R), brep);
        if constexpr (DELTA) {
            acc = swar_add<W>(acc, v);
            v   = acc;
        }
        V[p] = v;
    });
}

// Stores the untransposed segments of a 32-bit column's lanes. out points at the vector's first element.
template <int W>
__device__ __forceinline__ void store_segments32(const uint32_t (&V)[W], unsigned t, uint8_t* out) {
    if constexpr (W == 32) {
        uint4* o = reinterpret_cast<uint4*>(out + 4 * seg_start<32>(t));
        static_for<8>([&](auto q) { o[q] = make_uint4(V[4 * q], V[4 * q + 1], V[4 * q + 2], V[4 * q + 3]); });
    } else if constexpr (W == 16) {
        static_for<2>([&](auto l) {
            uint4* o = reinterpret_cast<uint4*>(out + 2 * seg_start<16>(2 * t + l));
            static_for<2>([&](auto q) {
                o[q] = make_uint4(gather_half<l>(V[8 * q + 0], V[8 * q + 1]), gather_half<l>(V[8 * q + 2], V[8 * q + 3]),
                                  gather_half<l>(V[8 * q + 4], V[8 * q + 5]), gather_half<l>(V[8 * q + 6], V[8 * q + 7]));
            });
        });
    } else {
        static_for<4>([&](auto l) {
            uint2* o = reinterpret_cast<uint2*>(out + seg_start<8>(4 * t + l));
            *o       = make_uint2(gather_byte<l>(V[0], V[1], V[2], V[3]), gather_byte<l>(V[4], V[5], V[6], V[7]));
        });
    }
}

// Lane-major ("transposed", the generated unffor's own output order) store: element j*L + i.
template <int W>
__device__ __forceinline__ void store_rows32_pmajor(const uint32_t (&V)[W], unsigned t, uint8_t* out) {
    uint32_t* o = reinterpret_cast<uint32_t*>(out);
    static_for<W>([&](auto p) {
        constexpr unsigned j = j_of_p<W>(p);
        __builtin_nontemporal_store(V[p], &o[j * 32 + t]);
    });
}

// W = 64: one lane per thread, values produced and stored in p order.
template <int BW, bool DELTA, bool LANE_MAJOR>
__device__ __forceinline__ void decode64(const uint8_t* packed, uint64_t base, uint64_t lane_base, unsigned t,
                                         uint8_t* out) {
    uint64_t R[BW > 0 ? BW : 1];
    load_rows64<BW>(packed, t, R);
    uint64_t acc = lane_base;
    uint64_t prev;
    static_for<64>([&](auto p) {
        constexpr unsigned j = j_of_p<64>(p);
        uint64_t           v = extract64<BW, j>(R) + base;
        if constexpr (DELTA) {
            acc += v;
            v = acc;
        }
        if constexpr (LANE_MAJOR) {
            __builtin_nontemporal_store(v, reinterpret_cast<uint64_t*>(out) + j * 16 + t);
        } else {
            if constexpr (p % 2 == 0) prev = v;
            else {
                uint4* o = reinterpret_cast<uint4*>(out + 8 * (64 * t + p - 1));
                *o       = make_uint4((uint32_t)prev, (uint32_t)(prev >> 32), (uint32_t)v, (uint32_t)(v >> 32));
            }
        }
    });
}

// 64-bit lane values in p order into an array (for gathers that need them all).
template <int BW, bool DELTA>
__device__ __forceinline__ void lanes_p64(const uint8_t* packed, uint64_t base, uint64_t lane_base, unsigned t,
                                          uint64_t (&V)[64]) {
    uint64_t R[BW > 0 ? BW : 1];
    load_rows64<BW>(packed, t, R);
    uint64_t acc = lane_base;
    static_for<64>([&](auto p) {
        uint64_t v = extract64<BW, j_of_p<64>(p)>(R) + base;
        if constexpr (DELTA) {
            acc += v;
            v = acc;
        }
        V[p] = v;
    });
}

// ---------------------------------------------------------------------------------------------------------------
// Gathers: output T-typed values out[seg + p] = table[index_p].
// ---------------------------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T load_t(const uint8_t* p) {
    if constexpr (sizeof(T) == 1) return *p;
    else if constexpr (sizeof(T) == 2) return *reinterpret_cast<const u16_ua*>(p);
    else if constexpr (sizeof(T) == 4) return *reinterpret_cast<const u32_ua*>(p);
    else return *reinterpret_cast<const u64_ua*>(p);
}

// Writes N consecutive T values from vals[] to out; ALIGN is the byte alignment guaranteed for out.
template <typename T, int N, int ALIGN>
__device__ __forceinline__ void store_run(const T (&vals)[N], T* out) {
    constexpr int bytes = ALIGN < 16 ? ALIGN : 16;
    constexpr int per   = bytes / (int)sizeof(T) > 0 ? bytes / (int)sizeof(T) : 1;
    if constexpr (per > 1 && N % per == 0 && bytes >= 4) {
        using V = typename std::conditional<bytes == 16, uint4, typename std::conditional<bytes == 8, uint2, uint32_t>::type>::type;
        static_for<N / per>([&](auto q) {
            union {
                T e[per];
                V u;
            } b;
            static_for<per>([&](auto k) { b.e[k] = vals[q * per + k]; });
            reinterpret_cast<V*>(out)[q] = b.u;
        });
    } else {
        static_for<N>([&](auto k) { out[k] = vals[k]; });
    }
}

}  // namespace fls
