// k_fsst.hip -- FSST string pages on gfx950.
//
// Page per vector (include/fls/cor/exp/fsst/fsst_exp.hpp:35-52, src/cor/lyt/vec.cpp:160-207): buffer 0 holds the
// 1024 compressed strings back to back, buffer 1 their cumulative end offsets (u32) in transposed order.
// Decoding follows include/fls/encoder/assert_eq.hpp:166-210: untranspose the offsets, then decode each string
// with fsst_decompress (include/fls/cor/prm/fsst/fsst.h:152-258): code < 255 emits symbol[code] (len[code] bytes),
// code 255 emits the next byte literally.
//
// One kernel, one pass over the compressed bytes. A block of 512 threads owns one half-vector segment (512
// strings). Its compressed bytes are split into equal 32-byte-aligned shares, one per thread, read straight from
// HBM into registers:
//   1. count: each thread adds up the decoded length of its share (a byte is a literal when it follows an odd run
//      of 255 bytes inside its string; an escape emits one byte, a code c emits len[c]);
//   2. one block scan of (bytes, string starts) gives every share its output position and the rank of its first
//      string start;
//   3. walk: each thread walks its share again, records the output position of every string start it passes, and
//      ORs its decoded bytes, 8 at a time, into a zeroed LDS image of the segment's output;
//   4. the segment's output base comes from a decoupled look-back over the column's segments (a status word per
//      segment: aggregate, then inclusive prefix; blocks take segments in dispatch order from a ticket counter so
//      every segment they wait on is already running), or from the caller's offsets;
//   5. the LDS image is copied out with coalesced 16-byte stores, and the string end offsets are written.
// Segments that are malformed (non-monotone offsets), larger than the LDS image, or whose compressed bytes exceed
// the start bitmap, take a per-string path that reads and writes global memory byte by byte.
#include "../../include/fls_gpu.h"
#include "gpu_internal.hpp"
#include "page_kernels.hpp"

namespace fls {

namespace {

constexpr unsigned kSeg      = 512;        // strings per segment (= threads per block)
constexpr unsigned kSegPerV  = 1024 / kSeg;
constexpr unsigned kWaves    = kSeg / 64;
constexpr uint32_t kMaxComp  = 32 * 1024;  // compressed bytes of a segment on the byte-parallel path
constexpr uint32_t kOutLds   = 24 * 1024;  // decoded bytes of a segment assembled in LDS
constexpr uint32_t kShare    = 32;         // share granule: one word of the start bitmap

constexpr unsigned kCallBase = 16;  // u64 index of the call state, 128 bytes past the ticket counter

// error bits (StringPlan::d_err)
constexpr uint32_t kErrCap = 4, kErrSpin = 16, kErrTicket = 32;

// Segment status word of the look-back: [63:48] call epoch, [47:46] flag, [45:0] decoded bytes.
constexpr uint64_t kFlagAgg = 1, kFlagPrefix = 2, kValMask = (1ull << 46) - 1;

struct SegCtl {
    uint64_t* status;       // one word per segment of the column
    uint64_t* ticket;       // [0] dispatch-order ticket counter (never reset); [kCallBase] its value when the call
                            // began, [kCallBase + 1] the call's epoch (1..65535), on a cache line of their own: set on
                            // the device by k_begin_call, so a captured graph of a call stays valid on replay
    uint64_t  ticket_off;   // tickets taken by the call's earlier page launches
    uint64_t  seg0;         // column-wide index of the page's first segment
    uint32_t  spin_limit;   // look-back polls before giving up (error bit kErrSpin)
};

__device__ __forceinline__ uint64_t status_word(uint32_t epoch, uint64_t flag, uint64_t v) {
    return ((uint64_t)epoch << 48) | (flag << 46) | (v & kValMask);
}

enum : int { FS_OFFSETS = 0, FS_BYTES = 1, FS_ALL = 2 };

// transposed index of original position i (inverse of untranspose_i): i = r + 8*ORDER[b] + 64c -> 128r + 16b + c
__device__ __forceinline__ unsigned transposed_of(unsigned i) {
    const unsigned c = i >> 6, b = order8((i >> 3) & 7), r = i & 7;
    return 128 * r + 16 * b + c;
}

// Exclusive scan of one u64 per thread over the block; returns the block total in *total.
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t x, uint64_t* total) {
    __shared__ uint64_t wsum[kWaves];
    const unsigned      lane = threadIdx.x % 64, w = threadIdx.x / 64;
    uint64_t            inc  = x;
    for (unsigned d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint64_t before = 0, all = 0;
    for (unsigned k = 0; k < kWaves; ++k) {
        const uint64_t s = wsum[k];
        if (k < w) before += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + inc - x;
}

// 32 bytes of a share into registers: two 16-byte loads, or byte loads for the last, partial granule of a segment
// (nothing at or past `avail` is read).
__device__ __forceinline__ void load_granule(const uint8_t* p, uint32_t avail, uint32_t (&wd)[8]) {
    if (avail >= kShare) {
        const v4u a = *reinterpret_cast<const v4u_ua*>(p), b = *reinterpret_cast<const v4u_ua*>(p + 16);
        wd[0] = a[0], wd[1] = a[1], wd[2] = a[2], wd[3] = a[3], wd[4] = b[0], wd[5] = b[1], wd[6] = b[2], wd[7] = b[3];
    } else {
        static_for<8>([&](auto d) { wd[d.value] = 0u; });
        static_for<32>([&](auto J) {
            constexpr unsigned j = decltype(J)::value;
            if (j < avail) wd[j / 4] |= (uint32_t)p[j] << (8 * (j % 4));
        });
    }
}

// Per-thread walk state over its share.
struct Walk {
    uint32_t lit;   // 1: the next byte is a literal consumed by an escape
    uint32_t run;   // decoded bytes so far (segment-relative in the second pass)
    uint32_t rk;    // rank of the next string start (second pass)
    uint32_t w;     // LDS output word
    uint32_t ab;    // pending bytes in acc
    uint64_t acc;   // bytes of word w not yet complete
};

// The granule passes below run as a loop of 4 rounds over 8 positions each. Between rounds the registers rotate:
// the round's bytes are always wd[0..1], the byte after them the low byte of wd[2], their start bits (and the next
// one) the low 9 bits of sb, their packed lengths kp[0]. Per-byte decisions are integer arithmetic (0/1 values), so
// no lane masks pile up across the unrolled steps.
__device__ __forceinline__ void rotate(uint32_t (&wd)[8], uint32_t (&kp)[4], uint64_t& sb, uint32_t fill) {
    static_for<6>([&](auto d) { wd[d.value] = wd[d.value + 2]; });
    wd[6] = fill, wd[7] = 0u;
    kp[0] = kp[1], kp[1] = kp[2], kp[2] = kp[3], kp[3] = 0u;
    sb >>= 8;
}

// Pass 1 over one granule (32 positions; `sb` holds their start bits and the next one; positions at or past `lim`
// are outside the share): decoded bytes into st.run and each position's length, in 4 bits, into kp. len[255] is 1
// (an escape emits one byte) and a literal emits nothing. Each round issues its 8 table loads together.
__device__ __forceinline__ void count_granule(const uint32_t (&wd_in)[8], uint64_t sb, uint32_t lim,
                                              const uint8_t* len, Walk& st, uint32_t (&kp)[4]) {
    uint32_t wd[8], unused[4] = {0u, 0u, 0u, 0u};
    static_for<8>([&](auto d) { wd[d.value] = wd_in[d.value]; });
#pragma unroll 1
    for (uint32_t r = 0; r < 4; ++r) {
        const int32_t lr = (int32_t)lim - (int32_t)(8 * r);
        uint32_t      lc[8];
        static_for<8>([&](auto J) { lc[J.value] = len[(wd[J.value / 4] >> (8 * (J.value % 4))) & 255u]; });
        uint32_t kw = 0;
        static_for<8>([&](auto J) {
            constexpr unsigned j   = decltype(J)::value;
            const uint32_t     c   = (wd[j / 4] >> (8 * (j % 4))) & 255u;
            const uint32_t     is  = (c + 1u) >> 8;                      // 1 for an escape byte
            const uint32_t     s1  = (uint32_t)(sb >> (j + 1)) & 1u;
            const uint32_t     in  = (uint32_t)((int32_t)j - lr) >> 31;  // j < lr
            const uint32_t     act = in & (st.lit ^ 1u);                 // a code or an escape inside the share
            const uint32_t     k   = lc[j] & (0u - act);
            st.lit                 = act & is & (s1 ^ 1u);
            st.run += k;
            kw |= k << (4 * j);
        });
        kp[0] = kp[1], kp[1] = kp[2], kp[2] = kp[3], kp[3] = kw;  // after 4 rounds kp[r] holds round r
        rotate(wd, unused, sb, 0u);
    }
}

// Pass 2 over one granule: records the output position of every string start (ostart by rank; a lane-private
// dummy slot takes the write where none starts) and, with EMIT, ORs the decoded bytes into obuf 8 at a time.
// Lengths come from pass 1 (HAVE_K: kp) or are recomputed with the literal chain. `nb` is the byte after the
// granule.
template <bool EMIT, bool HAVE_K>
__device__ __forceinline__ void walk_granule(const uint32_t (&wd_in)[8], uint32_t nb, uint64_t sb, uint32_t lim,
                                             const uint32_t (&kp_in)[4], const uint64_t* sym, const uint8_t* len,
                                             uint32_t* ostart, uint32_t dummy, uint64_t* obuf, Walk& st) {
    uint32_t wd[8], kp[4];
    static_for<8>([&](auto d) { wd[d.value] = wd_in[d.value]; });
    static_for<4>([&](auto d) { kp[d.value] = kp_in[d.value]; });
#pragma unroll 1
    for (uint32_t r = 0; r < 4; ++r) {
        const int32_t lr = (int32_t)lim - (int32_t)(8 * r);
        uint64_t      sv[8];
        uint32_t      lc[8];
        static_for<8>([&](auto J) {
            const uint32_t c = (wd[J.value / 4] >> (8 * (J.value % 4))) & 255u;
            if constexpr (EMIT) sv[J.value] = sym[c];  // sym[255] = 0; symbols are zero beyond their length
            if constexpr (!HAVE_K) lc[J.value] = len[c];
        });
        static_for<8>([&](auto J) {
            constexpr unsigned j  = decltype(J)::value;
            const uint32_t     c  = (wd[j / 4] >> (8 * (j % 4))) & 255u;
            const uint32_t     is = (c + 1u) >> 8;
            const uint32_t     s1 = (uint32_t)(sb >> (j + 1)) & 1u;
            const uint32_t     in = (uint32_t)((int32_t)j - lr) >> 31;
            uint32_t           k;
            if constexpr (HAVE_K) {
                k = (kp[0] >> (4 * j)) & 15u;
            } else {
                const uint32_t act = in & (st.lit ^ 1u);
                k                  = lc[j] & (0u - act);
                st.lit             = act & is & (s1 ^ 1u);
            }
            const uint32_t s0 = in & (uint32_t)(sb >> j) & 1u;
            ostart[dummy + s0 * (st.rk - dummy)] = st.run;
            st.rk += s0;
            if constexpr (EMIT) {
                const uint32_t nxt = j < 7 ? (wd[(j + 1) / 4] >> (8 * ((j + 1) % 4))) & 255u : wd[2] & 255u;
                // k != 0: a code (its symbol) or an escape (the next byte, or 0 when its string ends first)
                const uint64_t eb   = (uint64_t)(nxt * (is & (s1 ^ 1u)));
                const uint32_t kz   = (k + 15u) >> 4;  // 1 when k != 0
                const uint64_t keep = 0ull - (uint64_t)kz;
                const uint64_t x    = (sv[j] | eb) & keep;
                const uint32_t sh   = 8 * st.ab;
                const uint64_t lo   = st.acc | (x << sh);
                const uint64_t hi   = (x >> 1) >> (63 - sh);  // the bytes past the word (none when sh = 0)
                atomicOr(reinterpret_cast<unsigned long long*>(obuf + st.w), (unsigned long long)lo);
                const uint32_t all  = st.ab + k;
                const uint64_t full = (uint64_t)(all >> 3);
                st.w += (uint32_t)full;
                st.acc = lo ^ ((lo ^ hi) & (0ull - full));
                st.ab  = all & 7u;
                __builtin_amdgcn_sched_barrier(0);  // one byte at a time: interleaving steps only adds registers
            }
            st.run += k;
        });
        rotate(wd, kp, sb, nb);
        nb = 0u;
    }
}

// -- granules of codes only: no 0xFF byte, no literal at the start, all 32 positions inside the share ---------------
__device__ __forceinline__ bool has_ff(const uint32_t (&wd)[8]) {
    uint32_t z = 0;
    static_for<8>([&](auto d) {
        const uint32_t n = ~wd[d.value];
        z |= (n - 0x01010101u) & ~n & 0x80808080u;  // a zero byte of ~w: an 0xFF byte of w
    });
    return z != 0;
}

__device__ __forceinline__ uint32_t nib_sum(uint32_t v) {  // sum of the 8 nibbles of v (each <= 8)
    const uint32_t t = (v & 0x0F0F0F0Fu) + ((v >> 4) & 0x0F0F0F0Fu);
    return (t * 0x01010101u) >> 24;
}

// Pass 1: every byte is a code; its length goes to kp and the total to st.run.
__device__ __forceinline__ void count_codes(const uint32_t (&wd)[8], const uint8_t* len, Walk& st, uint32_t (&kp)[4]) {
    static_for<4>([&](auto R) {
        constexpr unsigned r = decltype(R)::value;
        uint32_t           lc[8];
        static_for<8>([&](auto J) { lc[J.value] = len[(wd[2 * r + J.value / 4] >> (8 * (J.value % 4))) & 255u]; });
        uint32_t kw = 0;
        static_for<8>([&](auto J) { kw |= lc[J.value] << (4 * J.value); });
        kp[r] = kw;
        st.run += nib_sum(kw);
    });
}

// Pass 2, bytes: OR each code's symbol into the LDS image (st.w / st.ab / st.acc as in walk_granule).
__device__ __forceinline__ void emit_codes(const uint32_t (&wd)[8], const uint32_t (&kp)[4], const uint64_t* sym,
                                           uint64_t* obuf, Walk& st) {
    static_for<4>([&](auto R) {
        constexpr unsigned r = decltype(R)::value;
        uint64_t           sv[8];
        static_for<8>([&](auto J) { sv[J.value] = sym[(wd[2 * r + J.value / 4] >> (8 * (J.value % 4))) & 255u]; });
        static_for<8>([&](auto J) {
            const uint32_t k  = (kp[r] >> (4 * J.value)) & 15u;
            const uint64_t x  = sv[J.value];
            const uint32_t sh = 8 * st.ab;
            const uint64_t lo = st.acc | (x << sh);
            const uint64_t hi = (x >> 1) >> (63 - sh);
            atomicOr(reinterpret_cast<unsigned long long*>(obuf + st.w), (unsigned long long)lo);
            const uint32_t all = st.ab + k;
            const bool     full = all >= 8;
            st.w += full;
            st.acc = full ? hi : lo;
            st.ab  = all & 7u;
            __builtin_amdgcn_sched_barrier(0);
        });
    });
}

// Pass 2, bytes, for any first granule (escapes, literals, positions past the share): lengths from pass 1 (a
// literal or a position past the share has length 0); a code emits its symbol, an escape the byte after it (0 when
// its string ends first).
__device__ __forceinline__ void emit_first(const uint32_t (&wd)[8], uint32_t nb, uint64_t sb, const uint32_t (&kp)[4],
                                           const uint64_t* sym, uint64_t* obuf, Walk& st) {
    static_for<4>([&](auto R) {
        constexpr unsigned r = decltype(R)::value;
        uint64_t           sv[8];
        static_for<8>([&](auto J) { sv[J.value] = sym[(wd[2 * r + J.value / 4] >> (8 * (J.value % 4))) & 255u]; });
        static_for<8>([&](auto J) {
            constexpr unsigned j   = 8 * r + decltype(J)::value;
            const uint32_t     c   = (wd[j / 4] >> (8 * (j % 4))) & 255u;
            const uint32_t     k   = (kp[r] >> (4 * J.value)) & 15u;
            const uint32_t     s1  = (uint32_t)(sb >> (j + 1)) & 1u;
            uint32_t           nxt;
            if constexpr (j + 1 < 32) nxt = (wd[(j + 1) / 4] >> (8 * ((j + 1) % 4))) & 255u;
            else nxt = nb;
            const uint64_t eb = (uint64_t)(nxt * (((c + 1u) >> 8) & (s1 ^ 1u)));
            const uint64_t x  = (sv[J.value] | eb) & (0ull - (uint64_t)((k + 15u) >> 4));
            const uint32_t sh = 8 * st.ab;
            const uint64_t lo = st.acc | (x << sh);
            const uint64_t hi = (x >> 1) >> (63 - sh);
            atomicOr(reinterpret_cast<unsigned long long*>(obuf + st.w), (unsigned long long)lo);
            const uint32_t all  = st.ab + k;
            const bool     full = all >= 8;
            st.w += full;
            st.acc = full ? hi : lo;
            st.ab  = all & 7u;
            __builtin_amdgcn_sched_barrier(0);
        });
    });
}

// Pass 2, positions: the output position of each string start among the granule's 32 positions (bits of sbm) is the
// granule's first position run0 plus the lengths before it (nibble sums of kp).
__device__ __forceinline__ void record_starts(uint32_t sbm, const uint32_t (&kp)[4], uint32_t run0, uint32_t* ostart,
                                              Walk& st) {
    const uint32_t s1 = nib_sum(kp[0]), s2 = s1 + nib_sum(kp[1]), s3 = s2 + nib_sum(kp[2]);
    while (sbm) {
        const uint32_t p = (uint32_t)__builtin_ctz(sbm), q = p >> 3;
        sbm &= sbm - 1;
        const uint32_t word   = q == 0 ? kp[0] : q == 1 ? kp[1] : q == 2 ? kp[2] : kp[3];
        const uint32_t before = q == 0 ? 0u : q == 1 ? s1 : q == 2 ? s2 : s3;
        ostart[st.rk++]       = run0 + before + nib_sum(word & ((1u << (4 * (p & 7))) - 1));
    }
    st.run = run0 + s3 + nib_sum(kp[3]);
}

// Wave 0 of the block: bytes decoded by all segments before g (decoupled look-back over their status words). Each
// poll reads a window of 512 predecessors (8 per lane), so the frontier of known prefixes advances 512 segments per
// memory round trip. Status words are read and written with relaxed agent-scope atomics: a block consumes only the
// words' own values (no data that other blocks wrote), and acquire/release at agent scope would invalidate or write
// back the whole L2 of the XCD at every poll.
__device__ __forceinline__ uint64_t look_back(const SegCtl& sc, uint32_t epoch, uint64_t g, uint32_t* err) {
    constexpr unsigned M     = 8;
    const unsigned     lane  = threadIdx.x % 64;
    uint64_t           excl  = 0;
    int64_t            j     = (int64_t)g - 1;
    uint32_t           spins = 0;
    while (j >= 0) {
        uint64_t val[M];
        uint32_t fl[M];
        static_for<M>([&](auto MM) {
            constexpr unsigned m   = decltype(MM)::value;
            const int64_t      idx = j - (int64_t)(lane + 64 * m);
            fl[m] = (uint32_t)kFlagPrefix, val[m] = 0;  // before the first segment: a prefix of 0
            if (idx >= 0) {
                const uint64_t w = __hip_atomic_load(sc.status + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                fl[m]            = (uint32_t)(w >> 48) == epoch ? (uint32_t)((w >> 46) & 3u) : 0u;
                val[m]           = w & kValMask;
            }
        });
        // distance of the nearest inclusive prefix; every segment nearer must have published its aggregate
        unsigned first = 64 * M;
        bool     wait  = false;
        static_for<M>([&](auto MM) {
            constexpr unsigned m  = decltype(MM)::value;
            const uint64_t     bp = __ballot(fl[m] == kFlagPrefix);
            const uint64_t     bx = __ballot(fl[m] == 0);
            if (first == 64 * M && !wait) {
                const unsigned f    = bp ? (unsigned)__builtin_ctzll(bp) : 64u;
                const uint64_t need = f >= 63 ? ~0ull : ((2ull << f) - 1);
                if (bx & need) wait = true;
                else if (bp) first = 64 * m + f;
            }
        });
        if (wait) {
            if (++spins > sc.spin_limit) {
                if (lane == 0) atomicOr(err, kErrSpin);
                return excl;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        uint64_t part = 0;
        static_for<M>([&](auto MM) {
            constexpr unsigned m = decltype(MM)::value;
            if (lane + 64 * m <= first) part += val[m];
        });
        for (unsigned d = 32; d >= 1; d >>= 1) part += __shfl_xor(part, d, 64);
        excl += part;
        if (first < 64 * M) break;
        j -= 64 * M;
    }
    return excl;
}

// One string decoded from global memory to global memory, byte by byte, into out[w, stop).
__device__ __forceinline__ void decode_string_global(const uint8_t* cb, uint32_t s, uint32_t e, const uint64_t* sym,
                                                     const uint8_t* len, uint8_t* out, uint64_t w, uint64_t stop) {
    for (uint32_t q = s; q < e && w < stop;) {
        const unsigned c = cb[q];
        if (c == 255) {
            out[w++] = q + 1 < e ? cb[q + 1] : 0;
            q += 2;
        } else {
            const uint64_t x = sym[c];
            const unsigned n = len[c];
            for (unsigned b = 0; b < n && w < stop; ++b) out[w++] = (uint8_t)(x >> (8 * b));
            q += 1;
        }
    }
}

// MODE FS_OFFSETS writes the n+1 string end offsets; FS_BYTES writes the bytes at the caller's offsets; FS_ALL
// writes both. `offsets` is indexed from the column's first string; `first_string` is the page's. The launch bound
// asks for 6 waves per SIMD (at most 80 VGPRs): three 512-thread blocks per compute unit, which the ~38 KB of LDS
// per block also allows; without it the compiler takes ~100 VGPRs and two blocks fit.
template <int MODE>
__global__ __launch_bounds__(kSeg, 6) void k_fsst(const uint8_t* __restrict__ page, const FsstTableDev* __restrict__ table,
                                               uint64_t first_string, uint64_t* __restrict__ offsets,
                                               uint8_t* __restrict__ out, uint64_t cap, uint32_t* err, SegCtl sc) {
    constexpr bool LOOKBACK = MODE != FS_BYTES;
    constexpr bool EMIT     = MODE != FS_OFFSETS;
    __shared__ uint64_t sym[256];
    __shared__ uint8_t  len[256];
    __shared__ uint32_t sbits[kMaxComp / 32 + 2];  // string starts over positions [0, n]
    __shared__ uint32_t ostart[kSeg + 1 + 64];     // output position by start rank; then one dummy slot per lane
    __shared__ uint16_t rank0[kSeg];               // start rank at each share's first position
    __shared__ uint64_t obuf[EMIT ? kOutLds / 8 + 1 : 1];
    __shared__ uint64_t bcast;
    const unsigned      tid = threadIdx.x, lane = tid % 64;

    uint64_t local = blockIdx.x;
    uint32_t epoch = 0;
    if constexpr (LOOKBACK) {
        // one thread reads the call's state and takes the ticket; the block shares them through LDS
        __shared__ uint32_t bepoch;
        if (tid == 0) {
            // written by k_begin_call, an earlier kernel of the stream: plain loads see it (and stay in the L2)
            const uint64_t base = sc.ticket[kCallBase];
            bepoch = (uint32_t)sc.ticket[kCallBase + 1];
            bcast  = __hip_atomic_fetch_add(sc.ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                    (base + sc.ticket_off);
        }
        __syncthreads();
        local = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(bcast >> 32)) << 32) |
                __builtin_amdgcn_readfirstlane((uint32_t)bcast);  // block-uniform: scalar registers
        epoch = __builtin_amdgcn_readfirstlane(bepoch);
        __syncthreads();
        if (local >= gridDim.x) {  // the host's ticket count disagrees with the counter: stop, report
            if (tid == 0) atomicOr(err, kErrTicket);
            return;
        }
    }
    __shared__ uint32_t ends_l[1024];
    const unsigned      v = (unsigned)(local / kSegPerV), h = (unsigned)(local % kSegPerV), i = h * kSeg + tid;
    const uint8_t*      cb   = vbuf(page, 2, v, 0);
    const uint8_t*      ends = vbuf(page, 2, v, 1);
    // the table, and the vector's 1024 (transposed) end offsets with one coalesced 16-byte load per thread
    if (tid < 256) {
        sym[tid] = table->symbol[tid], len[tid] = table->len[tid];
        *reinterpret_cast<v4u*>(ends_l + 4 * tid) = *reinterpret_cast<const v4u_ua*>(ends + 16 * tid);
    }
    __syncthreads();
    auto end_at = [&](int k) { return k < 0 ? 0u : ends_l[transposed_of((unsigned)k)]; };
    const uint64_t  g    = sc.seg0 + local;  // column-wide segment
    uint64_t* const oseg = offsets + first_string + (uint64_t)v * 1024 + h * kSeg;
    const uint32_t  s0 = __builtin_amdgcn_readfirstlane(end_at((int)(h * kSeg) - 1));
    const uint32_t  s1 = __builtin_amdgcn_readfirstlane(end_at((int)(h * kSeg + kSeg) - 1));
    uint32_t        s = end_at((int)i - 1), e = end_at((int)i);
    // byte-parallel shares: 32-byte granules, C bytes per thread
    const bool     span_ok = s0 <= s1 && s1 - s0 <= kMaxComp;  // block-uniform
    const uint32_t n       = span_ok ? s1 - s0 : 0;
    const uint32_t C       = kShare * ((n + kShare * kSeg - 1) / (kShare * kSeg));
    const uint32_t q0 = tid * C < n ? tid * C : n, q1 = q0 + C < n ? q0 + C : n;
    const uint8_t* seg = cb + s0;
    // the share's first granule, the byte after it and the byte before it, loaded ahead of the barriers below
    // (every byte lies in [s0, s1], inside the vector's compressed bytes)
    uint32_t wd0[8], nb0 = 0, prev = 0;
    if (q0 < q1) {
        load_granule(seg + q0, n - q0, wd0);
        nb0  = q0 + kShare < n ? seg[q0 + kShare] : 0u;
        prev = q0 ? seg[q0 - 1] : 0u;
    }
    const bool fast = __syncthreads_and(span_ok && s0 <= s && s <= e && e <= s1) != 0;

    uint64_t seg_total;
    uint32_t start_rel = 0, end_rel = 0;  // this thread's string, segment-relative output positions
    uint32_t n_starts = 0;
    Walk     st {};
    uint32_t kp0[4] = {0u, 0u, 0u, 0u}, lit1 = 0;  // first granule: lengths, literal after it
    bool     codes0 = false;                        // first granule holds codes only
    if (fast) {
        for (uint32_t k = tid; k <= n / 32 + 1; k += kSeg) sbits[k] = 0u;
        __syncthreads();
        atomicOr(&sbits[(s - s0) >> 5], 1u << ((s - s0) & 31));
        if (tid == 0) atomicOr(&sbits[n >> 5], 1u << (n & 31));
        __syncthreads();
        auto is_start = [&](uint32_t p) { return ((sbits[p >> 5] >> (p & 31)) & 1u) != 0; };
        // a share that starts inside a run of 255 bytes starts on a literal when the run before it is odd (a run
        // begins with an escape code: after a code or a literal comes a code)
        uint32_t lit0 = 0;
        if (q0 < q1 && !is_start(q0) && prev == 255u) {
            uint32_t p = q0 - 1;
            while (!is_start(p) && seg[p - 1] == 255) --p;
            lit0 = (q0 - p) & 1u;
        }
        // string starts in the share (bit q1 excluded: it belongs to the next share or is the end bit n)
        uint32_t ns = 0;
        for (uint32_t b = q0; b < q1; b += kShare) {
            const uint32_t m = q1 - b >= 32 ? ~0u : ((1u << (q1 - b)) - 1);
            ns += __popc(sbits[b >> 5] & m);
        }
        // pass 1: decoded bytes of the share
        st.lit = lit0;
        if (q0 < q1) {
            const uint64_t sb  = sbits[q0 >> 5] | ((uint64_t)sbits[(q0 >> 5) + 1] << 32);
            const uint32_t lim = q1 - q0 < kShare ? q1 - q0 : kShare;
            codes0             = lim == kShare && st.lit == 0 && !has_ff(wd0);
            // one form per wave: the codes-only one when every lane of the wave can take it
            if (__all(codes0)) count_codes(wd0, len, st, kp0);
            else count_granule(wd0, sb, lim, len, st, kp0);
            lit1 = st.lit;
        }
        for (uint32_t b = q0 + kShare; b < q1; b += kShare) {
            uint32_t wd[8], kp_unused[4];
            load_granule(seg + b, n - b, wd);
            const uint64_t sb  = sbits[b >> 5] | ((uint64_t)sbits[(b >> 5) + 1] << 32);
            const uint32_t lim = q1 - b < kShare ? q1 - b : kShare;
            count_granule(wd, sb, lim, len, st, kp_unused);
        }
        uint64_t       all;
        const uint64_t before = block_exclusive_scan((uint64_t)st.run | ((uint64_t)ns << 32), &all);
        seg_total              = (uint32_t)all;
        n_starts               = (uint32_t)(all >> 32);
        rank0[tid]             = (uint16_t)(before >> 32);
        if (tid == 0) ostart[n_starts] = (uint32_t)seg_total;
        st = Walk {lit0, (uint32_t)before, (uint32_t)(before >> 32), (uint32_t)before >> 3, (uint32_t)before & 7u, 0ull};
    } else {
        // per-string lengths from global memory (offsets clamped so that every read stays inside the vector)
        if (s > e) s = e;
        uint32_t tot = 0;
        for (uint32_t q = s; q < e;) {
            const unsigned c = cb[q];
            if (c == 255) tot += 1, q += 2;
            else tot += len[c], q += 1;
        }
        uint64_t       all;
        const uint64_t before = block_exclusive_scan(tot, &all);
        seg_total             = all;
        start_rel             = (uint32_t)before;
        end_rel               = (uint32_t)before + tot;
    }

    // publish this segment's aggregate early, so that later segments can sum past it while it decodes
    if constexpr (LOOKBACK) {
        if (tid == 0)
            __hip_atomic_store(sc.status + g, status_word(epoch, g == 0 ? kFlagPrefix : kFlagAgg, seg_total),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    uint64_t seg_base = 0;
    bool     given_ok = true;
    if constexpr (!LOOKBACK) {
        seg_base = oseg[0];
        given_ok = oseg[kSeg] >= seg_base && oseg[kSeg] - seg_base == seg_total;
    }
    const bool in_lds = EMIT && fast && given_ok && seg_total <= kOutLds;
    if (in_lds) {  // the LDS image starts zeroed: pass 2 ORs bytes into it
        for (uint32_t k = tid; k <= (uint32_t)seg_total / 8; k += kSeg) obuf[k] = 0ull;
        __syncthreads();
    }

    if (fast) {
        // pass 2: string-start positions (and, in LDS, the decoded bytes)
        const uint32_t dummy = kSeg + 1 + lane;
        if (q0 < q1) {  // the first granule from registers, with its lengths from pass 1
            const uint64_t sb  = sbits[q0 >> 5] | ((uint64_t)sbits[(q0 >> 5) + 1] << 32);
            const uint32_t lim = q1 - q0 < kShare ? q1 - q0 : kShare;
            const uint32_t run0 = st.run;
            if (in_lds) {
                if (__all(codes0)) emit_codes(wd0, kp0, sym, obuf, st);
                else emit_first(wd0, nb0, sb, kp0, sym, obuf, st);
            }
            record_starts((uint32_t)sb & (lim == kShare ? ~0u : (1u << lim) - 1), kp0, run0, ostart, st);
            st.lit = lit1;  // the literal chain continues into the next granule
        }
        for (uint32_t b = q0 + kShare; b < q1; b += kShare) {
            const uint64_t sb  = sbits[b >> 5] | ((uint64_t)sbits[(b >> 5) + 1] << 32);
            const uint32_t lim = q1 - b < kShare ? q1 - b : kShare;
            uint32_t       wd[8];
            load_granule(seg + b, n - b, wd);
            const uint32_t nb = b + kShare < n ? seg[b + kShare] : 0u;
            if (in_lds) walk_granule<true, false>(wd, nb, sb, lim, kp0, sym, len, ostart, dummy, obuf, st);
            else walk_granule<false, false>(wd, nb, sb, lim, kp0, sym, len, ostart, dummy, obuf, st);
        }
        if (in_lds && q0 < q1)
            atomicOr(reinterpret_cast<unsigned long long*>(obuf + st.w), (unsigned long long)st.acc);
        __syncthreads();
        // this thread's string: output positions of its start and end from their start ranks
        auto rank_of = [&](uint32_t p) -> uint32_t {
            if (p >= n) return n_starts;
            const uint32_t T = C == kShare ? p / kShare : p / C;  // one granule per share: a shift
            uint32_t       r = rank0[T];
            for (uint32_t w = T * C / 32; w < p / 32; ++w) r += __popc(sbits[w]);
            return r + __popc(sbits[p >> 5] & ((1u << (p & 31)) - 1));
        };
        start_rel = ostart[rank_of(s - s0)];
        end_rel   = ostart[rank_of(e - s0)];
    }

    // the segment's output base: wave 0 looks back after pass 2 (an earlier look-back finds more predecessors that
    // have not published yet and polls longer) and publishes the inclusive prefix
    if constexpr (LOOKBACK) {
        if (tid < 64) {
            const uint64_t excl = g == 0 ? 0ull : look_back(sc, epoch, g, err);
            if (tid == 0) {
                if (g != 0)
                    __hip_atomic_store(sc.status + g, status_word(epoch, kFlagPrefix, excl + seg_total),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                bcast = excl;
            }
        }
        __syncthreads();
        seg_base = bcast;
        oseg[tid + 1] = seg_base + end_rel;
        if (g == 0 && tid == 0) oseg[0] = 0;
    }

    if constexpr (EMIT) {
        if (in_lds) {
            if (seg_base + seg_total > cap) {
                if (tid == 0) atomicOr(err, kErrCap);
                return;
            }
            // copy-out: 16 bytes per thread per step (unaligned global stores), then the tail byte by byte
            const uint32_t nb = (uint32_t)seg_total, n16 = nb & ~15u;
            const uint8_t* src = reinterpret_cast<const uint8_t*>(obuf);
            uint8_t*       dst = out + seg_base;
            for (uint32_t k = 16 * tid; k < n16; k += 16 * kSeg)
                *reinterpret_cast<v4u_ua*>(dst + k) = *reinterpret_cast<const v4u*>(src + k);
            for (uint32_t k = n16 + tid; k < nb; k += kSeg) dst[k] = src[k];
        } else {
            uint64_t w0, stop;
            if (LOOKBACK || given_ok) w0 = seg_base + start_rel, stop = seg_base + end_rel;
            else w0 = oseg[tid], stop = oseg[tid + 1];  // the caller's offsets, string by string
            if (s > e) s = e;
            if (stop > cap || stop < w0) atomicOr(err, kErrCap);
            else decode_string_global(cb, s, e, sym, len, out, w0, stop);
        }
    }
}

__global__ void k_zero_first(uint64_t* offsets) { offsets[0] = 0; }

// Start of a look-back call: the ticket counter's current value becomes the call's base, and the epoch advances.
__global__ void k_begin_call(uint64_t* ticket) {
    ticket[kCallBase]     = __hip_atomic_load(ticket, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ticket[kCallBase + 1] = ticket[kCallBase + 1] % 0xffffu + 1;
}

template <int MODE>
int launch_pages(const Column& col, StringPlan& sp, const uint8_t* dev_enc, uint64_t* dev_offsets, uint8_t* dev_bytes,
                 uint64_t cap, hipStream_t s) {
    if (MODE != FS_BYTES) hipLaunchKernelGGL(k_begin_call, dim3(1), dim3(1), 0, s, sp.d_ticket);
    uint64_t ticket_off = 0;
    for (size_t i = 0; i < col.pages.size(); ++i) {
        const PageEntry& pe = col.pages[i];
        if (!pe.vec_c) continue;
        const unsigned grid = pe.vec_c * kSegPerV;
        SegCtl         sc {sp.d_status, sp.d_ticket, ticket_off, (pe.out_index / 1024) * kSegPerV, 1u << 20};
        hipLaunchKernelGGL(k_fsst<MODE>, dim3(grid), dim3(kSeg), 0, s, dev_enc + pe.page_off, sp.d_tables + i,
                           pe.out_index, dev_offsets, dev_bytes, cap, sp.d_err, sc);
        ticket_off += grid;
    }
    if (MODE != FS_BYTES && col.n_values == 0) {
        // offsets[0] of an empty column, by a kernel: a memset node recorded into a caller's captured graph is not
        // reliable on replay
        hipLaunchKernelGGL(k_zero_first, dim3(1), dim3(1), 0, s, dev_offsets);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "kernel launch");
}

}  // namespace

int strings_prepare(const Column& col, const uint8_t* host, StringPlan& sp) {
    (void)host;
    std::vector<FsstTableDev> tabs(col.pages.size());
    for (size_t i = 0; i < col.pages.size(); ++i) {
        std::memset(&tabs[i], 0, sizeof(FsstTableDev));
        for (int c = 0; c < 255; ++c) {
            // bytes beyond a symbol's length must be zero: the decoder ORs whole symbols into its output words
            const unsigned l = col.pages[i].fsst.len[c];
            uint64_t       x = col.pages[i].fsst.symbol[c];
            if (l < 8) x &= (1ull << (8 * l)) - 1;
            tabs[i].symbol[c] = x, tabs[i].len[c] = (uint8_t)l;
        }
        tabs[i].len[255] = 1;
    }
    sp.n_vec           = col.n_values / 1024;
    const uint64_t nsg = sp.n_vec * kSegPerV;
    hipError_t     e;
    if ((e = hipMalloc(&sp.d_tables, tabs.size() * sizeof(FsstTableDev) + 1)) != hipSuccess) return hip_fail(e, "hipMalloc");
    if ((e = hipMemcpy(sp.d_tables, tabs.data(), tabs.size() * sizeof(FsstTableDev), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "hipMemcpy");
    if ((e = hipMalloc(&sp.d_status, (nsg + 1) * sizeof(uint64_t))) != hipSuccess) return hip_fail(e, "hipMalloc");
    if ((e = hipMemset(sp.d_status, 0, (nsg + 1) * sizeof(uint64_t))) != hipSuccess) return hip_fail(e, "hipMemset");
    if ((e = hipMalloc(&sp.d_ticket, (kCallBase + 2) * sizeof(uint64_t))) != hipSuccess) return hip_fail(e, "hipMalloc");
    if ((e = hipMemset(sp.d_ticket, 0, (kCallBase + 2) * sizeof(uint64_t))) != hipSuccess) return hip_fail(e, "hipMemset");
    if ((e = hipMalloc(&sp.d_err, 4)) != hipSuccess) return hip_fail(e, "hipMalloc");
    if ((e = hipMemset(sp.d_err, 0, 4)) != hipSuccess) return hip_fail(e, "hipMemset");
    return FLS_OK;
}

void strings_release(StringPlan& sp) {
    if (sp.d_tables) (void)hipFree(sp.d_tables);
    if (sp.d_status) (void)hipFree(sp.d_status);
    if (sp.d_ticket) (void)hipFree(sp.d_ticket);
    if (sp.d_err) (void)hipFree(sp.d_err);
    sp = StringPlan {};
}

int strings_offsets(const Column& col, StringPlan& sp, const uint8_t* dev_enc, uint64_t* dev_offsets, hipStream_t s) {
    return launch_pages<FS_OFFSETS>(col, sp, dev_enc, dev_offsets, nullptr, 0, s);
}

int strings_decode(const Column& col, StringPlan& sp, const uint8_t* dev_enc, const uint64_t* dev_offsets,
                   uint8_t* dev_bytes, uint64_t cap, hipStream_t s) {
    return launch_pages<FS_BYTES>(col, sp, dev_enc, const_cast<uint64_t*>(dev_offsets), dev_bytes, cap, s);
}

int strings_decode_all(const Column& col, StringPlan& sp, const uint8_t* dev_enc, uint64_t* dev_offsets,
                       uint8_t* dev_bytes, uint64_t cap, hipStream_t s) {
    return launch_pages<FS_ALL>(col, sp, dev_enc, dev_offsets, dev_bytes, cap, s);
}

int strings_check(StringPlan& sp, hipStream_t s) {
    uint32_t   h = 0;
    hipError_t e = hipMemcpyAsync(&h, sp.d_err, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "check");
    if (h & (kErrSpin | kErrTicket)) return fail(FLS_E_DATA, "string offsets: segment look-back did not complete");
    return h ? fail(FLS_E_DATA, "string offsets exceed the output capacity") : FLS_OK;
}

}  // namespace fls
