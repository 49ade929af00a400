// fls_lanes.hpp -- FastLanes 1024-value vector layout, as device code for gfx950.
//
// Layout facts (from the reference's generated kernels, primitives/fls_generated/fallback/scalar_aav_uf1/
// fallback_scalar_aav_1024_uf1_{unffor,rsum,untranspose}_src.cpp):
//   * W-bit elements, L = 1024/W lanes. Packed word w of lane i is in[w*L + i]; value j of lane i is at
//     transposed index j*L + i. Inside a lane, bw-bit values are concatenated little-endian.
//   * Transposed index k = 128r + 16b + c holds original index r + 8*ORDER[b] + 64c, ORDER = {0,4,2,6,1,5,3,7}.
//   * Consequence used throughout: lane i's W values occupy a contiguous run of W original positions,
//       W=8 : start 64(i%16) + 8*ORDER[i/16], position p = j
//       W=16: start 64(i%16) + 8*ORDER[i/16], position p = j/2 + 8(j%2)
//       W=32: start 64(i%16) + 32(i/16),      position p = 8*ORD4[j%4] + j/4     (ORD4 = {0,2,1,3})
//       W=64: start 64i,                       position p = 8*ORDER[j%8] + j/8
//     and the delta (rsum) scan of lane i visits exactly p = 0, 1, ..., W-1 in order.
//   So unpack, FFOR, delta and untranspose all happen in one thread's registers: no cross-lane traffic.
//
// Thread mapping: for W <= 32 a thread owns one 32-bit column of every packed row (G = 32/W adjacent lanes,
// 32 threads per vector, SWAR arithmetic inside the dword); for W = 64 a thread owns one 64-bit lane
// (16 threads per vector).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fls {

typedef uint32_t __attribute__((aligned(1))) u32_ua;
typedef uint64_t __attribute__((aligned(1))) u64_ua;
typedef uint16_t __attribute__((aligned(1))) u16_ua;

__host__ __device__ constexpr unsigned order8(unsigned b) {
    return (b & 1) * 4 + (b & 2) + (b >> 2);  // bit reversal of 3 bits: {0,4,2,6,1,5,3,7}
}
__host__ __device__ constexpr unsigned ord4(unsigned m) { return ((m & 1) << 1) | (m >> 1); }  // {0,2,1,3}

// j (value index within a lane) of untransposed position p within the lane's segment.
template <int W>
__host__ __device__ constexpr unsigned j_of_p(unsigned p) {
    if constexpr (W == 8) return p;
    else if constexpr (W == 16) return 2 * (p % 8) + p / 8;
    else if constexpr (W == 32) return 4 * (p % 8) + ord4(p / 8);
    else return 8 * (p % 8) + order8(p / 8);
}
// Start of lane i's segment in the untransposed vector.
template <int W>
__host__ __device__ constexpr unsigned seg_start(unsigned i) {
    if constexpr (W == 8 || W == 16) return 64 * (i % 16) + 8 * order8(i / 16);
    else if constexpr (W == 32) return 64 * (i % 16) + 32 * (i / 16);
    else return 64 * i;
}

// Per-lane replication of a W-bit constant inside a 32-bit dword (W <= 32).
template <int W>
__host__ __device__ constexpr uint32_t rep(uint32_t m) {
    if constexpr (W == 8) return (m & 0xffu) * 0x01010101u;
    else if constexpr (W == 16) return (m & 0xffffu) * 0x00010001u;
    else return m;
}
template <int W>
__host__ __device__ constexpr uint64_t lowmask(unsigned bits) {
    return bits >= 64 ? ~0ull : ((1ull << bits) - 1);
}

// Lane-wise add modulo 2^W inside a dword (W <= 32).
template <int W>
__device__ __forceinline__ uint32_t swar_add(uint32_t a, uint32_t b) {
    if constexpr (W == 32) return a + b;
    else {
        constexpr uint32_t H = W == 8 ? 0x80808080u : 0x80008000u;
        return ((a & ~H) + (b & ~H)) ^ ((a ^ b) & H);
    }
}

// Value j of every lane of a packed 32-bit column (W <= 32), BW bits per value, rows R[0..BW-1].
template <int W, int BW, int J>
__device__ __forceinline__ uint32_t extract32(const uint32_t* R) {
    if constexpr (BW == 0) return 0u;
    else if constexpr (BW == W) return R[J];
    else {
        constexpr unsigned bit = J * BW, w = bit / W, s = bit % W;
        if constexpr (s + BW <= W) {
            return (R[w] >> s) & rep<W>((uint32_t)lowmask<W>(BW));
        } else {
            constexpr unsigned lo_bits = W - s;
            constexpr uint32_t mlo     = rep<W>((uint32_t)lowmask<W>(lo_bits));
            constexpr uint32_t mhi     = rep<W>((uint32_t)(lowmask<W>(BW) & ~lowmask<W>(lo_bits)));
            return ((R[w] >> s) & mlo) | ((R[w + 1] << lo_bits) & mhi);
        }
    }
}
// Value j of a 64-bit lane.
template <int BW, int J>
__device__ __forceinline__ uint64_t extract64(const uint64_t* R) {
    if constexpr (BW == 0) return 0ull;
    else if constexpr (BW == 64) return R[J];
    else {
        constexpr unsigned bit = J * BW, w = bit / 64, s = bit % 64;
        if constexpr (s + BW <= 64) {
            return (R[w] >> s) & lowmask<64>(BW);
        } else {
            return ((R[w] >> s) | (R[w + 1] << (64 - s))) & lowmask<64>(BW);
        }
    }
}

template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    [&]<int... I>(std::integer_sequence<int, I...>) { (f(std::integral_constant<int, I> {}), ...); }
    (std::make_integer_sequence<int, N> {});
}

// Byte l (0..3) of each of four dwords, packed into one dword (a, b, c, d -> bytes 0..3).
template <int L>
__device__ __forceinline__ uint32_t gather_byte(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    // v_perm_b32 selects bytes of {src0, src1}: selector byte k picks byte sel of the 64-bit {src0:src1}.
    const uint32_t ab = __builtin_amdgcn_perm(b, a, (L) | ((L + 4) << 8) | (L << 16) | ((L + 4) << 24));
    const uint32_t cd = __builtin_amdgcn_perm(d, c, (L) | ((L + 4) << 8) | (L << 16) | ((L + 4) << 24));
    return __builtin_amdgcn_perm(cd, ab, 0x05040100u);
}
// 16-bit half l (0/1) of two dwords, packed (a -> low half, b -> high half).
template <int L>
__device__ __forceinline__ uint32_t gather_half(uint32_t a, uint32_t b) {
    constexpr uint32_t sel = L == 0 ? 0x05040100u : 0x07060302u;
    return __builtin_amdgcn_perm(b, a, sel);
}

}  // namespace fls
