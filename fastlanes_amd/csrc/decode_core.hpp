// decode_core.hpp -- per-vector decode bodies (FFOR / unpack, delta, RLE gather, dictionary gather) for gfx950.
//
// Two thread mappings, one per output order; every body is fully unrolled for a compile-time bit width BW.
//  * Lane-major output (the reference unffor's own order, raw streams): every packed row and every value row of a
//    1024-value vector is 128 bytes, so 8 threads decode a vector, thread q owning bytes [16q, 16q+16) of every
//    row (4 lanes for W=32, 2 for W=64, 8 / 16 SWAR lanes for W=16 / W=8). Rows are consumed in order, so few
//    registers are live; loads and stores are 16-byte.
//  * Untransposed output (pages): thread t owns one 32-bit column of every row (G = 32/W lanes, 32 threads per
//    vector) or one 64-bit lane (W = 64, 16 threads per vector). Lane i's values fill a contiguous segment of the
//    vector in position order p (fls_lanes.hpp), which is also the delta scan order, and are written with 16-byte
//    stores. A 16-byte column would have to hold 128 values in position order, which costs too many VGPRs.
// Reference semantics followed:
//   unffor / unpack  primitives/fls_generated/fallback/scalar_aav_uf1/fallback_scalar_aav_1024_uf1_unffor_src.cpp,
//                    ..._unpack_src.cpp (value j of lane i = bits [j*bw, j*bw+bw) of the lane's words, + base)
//   rsum             ..._rsum_src.cpp (lane base + running sum in the lane's visiting order, mod 2^W)
//   untranspose_i    ..._untranspose_src.cpp, src/encoder/materializer.cpp:13-21
//   RLE gather       src/cor/lyt/vec.cpp:318-332;  dictionary gather  include/fls/encoder/assert_eq.hpp:257-328
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

typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef uint32_t v4u_ua __attribute__((ext_vector_type(4), aligned(1)));
typedef uint32_t v2u __attribute__((ext_vector_type(2)));

constexpr unsigned kTPV = 8;  // threads per vector, lane-major mapping

// ---------------------------------------------------------------------------------------------------------------
// Memory helpers. Page buffers carry no alignment guarantee (a page is a byte stream), so packed rows are read
// through alignment-1 vector types; gfx950 executes these as ordinary dwordx4 loads.
// ---------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ v4u ld16(const uint8_t* p) { return __builtin_nontemporal_load(reinterpret_cast<const v4u_ua*>(p)); }

template <bool NT>
__device__ __forceinline__ void st16(uint8_t* p, v4u x) {
    if constexpr (NT) __builtin_nontemporal_store(x, reinterpret_cast<v4u*>(p));
    else *reinterpret_cast<v4u*>(p) = x;
}
template <bool NT>
__device__ __forceinline__ void st8(uint8_t* p, v2u x) {
    if constexpr (NT) __builtin_nontemporal_store(x, reinterpret_cast<v2u*>(p));
    else *reinterpret_cast<v2u*>(p) = x;
}

// Output sinks for the lane-column (page) bodies: byte offsets are relative to the vector's first value.
// LdsSink stages a vector in LDS with 16 bytes of padding per 256 (conflict-free 16-byte lane writes); the kernel
// then writes whole lines to HBM. GlobalSink writes in place.
__device__ __forceinline__ uint32_t lds_pad(uint32_t off) { return off + (off >> 8) * 16; }
struct LdsSink {
    uint8_t* base;
    __device__ __forceinline__ void put16(uint32_t off, v4u x) const { *reinterpret_cast<v4u*>(base + lds_pad(off)) = x; }
    __device__ __forceinline__ void put8(uint32_t off, v2u x) const { *reinterpret_cast<v2u*>(base + lds_pad(off)) = x; }
    template <typename T>
    __device__ __forceinline__ void put(uint32_t off, T x) const { *reinterpret_cast<T*>(base + lds_pad(off)) = x; }
};
template <bool NT>
struct GlobalSink {
    uint8_t* base;
    __device__ __forceinline__ void put16(uint32_t off, v4u x) const { st16<NT>(base + off, x); }
    __device__ __forceinline__ void put8(uint32_t off, v2u x) const { st8<NT>(base + off, x); }
    template <typename T>
    __device__ __forceinline__ void put(uint32_t off, T x) const { *reinterpret_cast<T*>(base + off) = x; }
};

template <int BW>
__device__ __forceinline__ void load_rows(const uint8_t* packed, unsigned q, v4u (&R)[BW > 0 ? BW : 1]) {
    static_for<BW>([&](auto w) { R[w] = ld16(packed + 128 * w + 16 * q); });
}

// Value j of the lanes in dword D of the thread's column (W <= 32).
template <int W, int BW, int J, int D>
__device__ __forceinline__ uint32_t value32(const v4u (&R)[BW > 0 ? BW : 1]) {
    if constexpr (BW == 0) return 0u;
    else if constexpr (BW == W) return R[J][D];
    else {
        constexpr unsigned bit = J * BW, w = bit / W, s = bit % W;
        if constexpr (s + BW <= W) {
            return (R[w][D] >> s) & rep<W>((uint32_t)lowmask<W>(BW));
        } else {
            constexpr unsigned lo_bits = W - s;
            constexpr uint32_t mlo     = rep<W>((uint32_t)lowmask<W>(lo_bits));
            constexpr uint32_t mhi     = rep<W>((uint32_t)(lowmask<W>(BW) & ~lowmask<W>(lo_bits)));
            return ((R[w][D] >> s) & mlo) | ((R[w + 1][D] << lo_bits) & mhi);
        }
    }
}
// Value j of 64-bit lane L (0/1) of the thread's column.
template <int BW, int J, int L>
__device__ __forceinline__ uint64_t value64(const v4u (&R)[BW > 0 ? BW : 1]) {
    auto word = [&](int w) { return (uint64_t)R[w][2 * L] | ((uint64_t)R[w][2 * L + 1] << 32); };
    if constexpr (BW == 0) return 0ull;
    else if constexpr (BW == 64) return word(J);
    else {
        constexpr unsigned bit = J * BW, w = bit / 64, s = bit % 64;
        if constexpr (s + BW <= 64) return (word(w) >> s) & lowmask<64>(BW);
        else return ((word(w) >> s) | (word(w + 1) << (64 - s))) & lowmask<64>(BW);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Lane-major output (the reference unffor/unpack's own order): row j = value j of all lanes, 128 bytes.
// ---------------------------------------------------------------------------------------------------------------
template <int W, int BW, bool NT>
__device__ __forceinline__ void ffor_lane_major(const uint8_t* packed, uint64_t base, unsigned q, uint8_t* out) {
    v4u R[BW > 0 ? BW : 1];
    load_rows<BW>(packed, q, R);
    if constexpr (W == 64) {
        static_for<64>([&](auto j) {
            const uint64_t a = value64<BW, j, 0>(R) + base, b = value64<BW, j, 1>(R) + base;
            st16<NT>(out + 128 * j + 16 * q, v4u {(uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32)});
        });
    } else {
        const uint32_t brep = rep<W>((uint32_t)base);
        static_for<W>([&](auto j) {
            v4u o;
            static_for<4>([&](auto d) { o[d.value] = swar_add<W>(value32<W, BW, j, d>(R), brep); });
            st16<NT>(out + 128 * j + 16 * q, o);
        });
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Lane-column mapping: row loads of one 32-bit column / 64-bit lane per thread.
// ---------------------------------------------------------------------------------------------------------------
template <int BW>
__device__ __forceinline__ void lane_rows32(const uint8_t* packed, unsigned t, uint32_t (&R)[BW > 0 ? BW : 1]) {
    const u32_ua* p = reinterpret_cast<const u32_ua*>(packed);
    static_for<BW>([&](auto w) { R[w] = __builtin_nontemporal_load(&p[w * 32 + t]); });
}
template <int BW>
__device__ __forceinline__ void lane_rows64(const uint8_t* packed, unsigned t, uint64_t (&R)[BW > 0 ? BW : 1]) {
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
    lane_rows32<BW>(packed, t, R);
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
template <int W, typename S>
__device__ __forceinline__ void store_segments32(const uint32_t (&V)[W], unsigned t, const S& out) {
    if constexpr (W == 32) {
        const uint32_t o = 4 * seg_start<32>(t);
        static_for<8>([&](auto q) { out.put16(o + 16 * q, v4u {V[4 * q], V[4 * q + 1], V[4 * q + 2], V[4 * q + 3]}); });
    } else if constexpr (W == 16) {
        static_for<2>([&](auto l) {
            const uint32_t o = 2 * seg_start<16>(2 * t + l);
            static_for<2>([&](auto q) {
                out.put16(o + 16 * q, v4u {gather_half<l>(V[8 * q + 0], V[8 * q + 1]), gather_half<l>(V[8 * q + 2], V[8 * q + 3]),
                                           gather_half<l>(V[8 * q + 4], V[8 * q + 5]), gather_half<l>(V[8 * q + 6], V[8 * q + 7])});
            });
        });
    } else {
        static_for<4>([&](auto l) {
            out.put8(seg_start<8>(4 * t + l), v2u {gather_byte<l>(V[0], V[1], V[2], V[3]), gather_byte<l>(V[4], V[5], V[6], V[7])});
        });
    }
}

// W = 64: one lane per thread, values produced and stored in p order.
template <int BW, bool DELTA, typename S>
__device__ __forceinline__ void decode64(const uint8_t* packed, uint64_t base, uint64_t lane_base, unsigned t,
                                         const S& out) {
    uint64_t R[BW > 0 ? BW : 1];
    lane_rows64<BW>(packed, t, R);
    uint64_t acc = lane_base;
    uint64_t prev;
    static_for<64>([&](auto p) {
        uint64_t v = extract64<BW, j_of_p<64>(p)>(R) + base;
        if constexpr (DELTA) {
            acc += v;
            v = acc;
        }
        if constexpr (p % 2 == 0) prev = v;
        else out.put16(8 * (64 * t + p - 1), v4u {(uint32_t)prev, (uint32_t)(prev >> 32), (uint32_t)v, (uint32_t)(v >> 32)});
    });
}

// ---------------------------------------------------------------------------------------------------------------
// Gathers
// ---------------------------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T load_t(const uint8_t* p) {
    if constexpr (sizeof(T) == 1) return *p;
    else if constexpr (sizeof(T) == 2) return *reinterpret_cast<const u16_ua*>(p);
    else if constexpr (sizeof(T) == 4) return *reinterpret_cast<const u32_ua*>(p);
    else return *reinterpret_cast<const u64_ua*>(p);
}

// Writes N consecutive T values at byte offset off of the sink; ALIGN is the alignment guaranteed for off.
template <typename T, int N, int ALIGN, typename S>
__device__ __forceinline__ void store_run(const T (&vals)[N], uint32_t off, const S& out) {
    constexpr int bytes = ALIGN < 16 ? ALIGN : 16;
    constexpr int per   = bytes / (int)sizeof(T);
    if constexpr (bytes == 16 && (N * sizeof(T)) % 16 == 0) {
        static_for<N / per>([&](auto k) {
            union {
                T   e[per];
                v4u u;
            } b;
            static_for<per>([&](auto e) { b.e[e] = vals[k * per + e]; });
            out.put16(off + 16 * k, b.u);
        });
    } else if constexpr (bytes == 8 && (N * sizeof(T)) % 8 == 0) {
        static_for<N / per>([&](auto k) {
            union {
                T   e[per];
                v2u u;
            } b;
            static_for<per>([&](auto e) { b.e[e] = vals[k * per + e]; });
            out.put8(off + 8 * k, b.u);
        });
    } else {
        static_for<N>([&](auto k) { out.template put<T>(off + sizeof(T) * k, vals[k]); });
    }
}

}  // namespace fls
