// encode.hip -- GPU encoder for integer columns: UNCOMPRESSED (pool id 0), FFOR (1), DELTA (11) and RLE (6, 7)
// pages, byte for byte what the host encoder (encode_from_memory) writes.
//
// Encode RPNs: UNCOMPRESSED {TRS, MEM_CPY} (include/fls/cor/exp/uncompressed/uncompressed_exp.hpp:32-47), FFOR
// {TRS, ANALYZE, FFOR} (include/fls/cor/exp/ffor/ffor_exp.hpp:39-46), DELTA {TRS, RSUM, ANALYZE, FFOR}
// (include/fls/cor/exp/delta/delta_exp.hpp:39-46); ANALYZE takes base = min and bw = the bit length of max - min
// (src/cor/prm/ffor/ffor_analyze.cpp:78-102, exact); pages as TCompressor::FlushPage / Page::Absorb
// (src/cor/eng/compressor.cpp:41-86, src/cor/lyt/page/page.cpp:87-130).
//
// Two passes: k_enc_analyze finds each vector's base and bit width (one wave per vector); the host reads the bit
// widths back, lays the pages out exactly as the host encoder does (65535 vectors and 128 MiB per page at most) and
// writes the page headers; k_enc_pack writes every vector's buffers and entry points (one block per vector: the
// vector is staged in LDS, thread i < L packs lane i).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/fls_gpu.h"
#include "decode_core.hpp"
#include "gpu_internal.hpp"

namespace fls {
namespace {

// original position of transposed index k (k = 128r + 16b + c -> r + 8 ORDER[b] + 64c)
__device__ __forceinline__ unsigned enc_orig(unsigned k) {
    const unsigned b = (k >> 4) & 7;
    return (k >> 7) + 8 * (((b & 1) << 2) | (b & 2) | (b >> 2)) + 64 * (k & 15);
}

// The value FFOR packs at original position o: the value itself, or (DELTA) its difference from the previous value
// inside the aligned W-value segment (0 at a segment start) -- the unrsum of the transposed lane order.
template <typename T, int MODE>
__device__ __forceinline__ T enc_value(const T* xs, unsigned o) {
    constexpr unsigned W = 8 * sizeof(T);
    if constexpr (MODE == 2) return (o % W) ? (T)(xs[o] - xs[o - 1]) : (T)0;
    else return xs[o];
}

template <typename T>
__device__ __forceinline__ T wave_min(T x) {
    for (int d = 32; d >= 1; d >>= 1) {
        const T y = __shfl_xor(x, d, 64);
        x         = y < x ? y : x;
    }
    return x;
}
template <typename T>
__device__ __forceinline__ T wave_max(T x) {
    for (int d = 32; d >= 1; d >>= 1) {
        const T y = __shfl_xor(x, d, 64);
        x         = y > x ? y : x;
    }
    return x;
}

// One wave per vector: base = min, bw = bit length of max - min over the values FFOR will pack. SGN: min and max
// of int64 columns compare as signed values (ffor_prm::ResolveAnalyzeFunc, src/cor/prm/ffor/ffor_analyze.cpp:145-160),
// done here on the values with their sign bit flipped.
template <typename T, int MODE, bool SGN>
__global__ __launch_bounds__(256) void k_enc_analyze(const T* __restrict__ x, uint64_t n_vec, uint8_t* __restrict__ bw,
                                                     uint64_t* __restrict__ base) {
    using C              = std::conditional_t<sizeof(T) == 8, unsigned long long, unsigned>;  // shuffle type
    constexpr C    bias  = SGN ? (C)((T)1 << (8 * sizeof(T) - 1)) : (C)0;
    const unsigned lane  = threadIdx.x % 64;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
    for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; v < n_vec; v += waves) {
        const T* xv = x + v * 1024;
        C        mn = (C)(T)~(T)0, mx = 0;
        for (unsigned o = lane; o < 1024; o += 64) {
            const C val = (C)enc_value<T, MODE>(xv, o) ^ bias;
            mn          = val < mn ? val : mn;
            mx          = val > mx ? val : mx;
        }
        mn = wave_min<C>(mn), mx = wave_max<C>(mx);
        if (lane == 0) {
            const uint64_t d = (uint64_t)(T)(mx - mn);
            bw[v]            = (uint8_t)(d ? 64 - __clzll((long long)d) : 0);
            base[v]          = (uint64_t)(T)(mn ^ bias);
        }
    }
}

struct EncPage {
    uint64_t start, v0;  // byte offset of the page in the output, its first vector
    uint32_t off[6];     // buffer offsets from the page start
};

template <typename T>
__device__ __forceinline__ void st_ua(uint8_t* p, T x) {
    if constexpr (sizeof(T) == 1) *p = x;
    else if constexpr (sizeof(T) == 2) *reinterpret_cast<u16_ua*>(p) = x;
    else if constexpr (sizeof(T) == 4) *reinterpret_cast<u32_ua*>(p) = x;
    else *reinterpret_cast<u64_ua*>(p) = x;
}

// One block of 128 threads per vector.
template <typename T, int MODE>
__global__ __launch_bounds__(128) void k_enc_pack(const T* __restrict__ x, const uint8_t* __restrict__ bw_in,
                                                  const uint64_t* __restrict__ base_in,
                                                  const uint64_t* __restrict__ pk_rel, const uint32_t* __restrict__ vpage,
                                                  const EncPage* __restrict__ pages, uint32_t arr_c,
                                                  uint8_t* __restrict__ out) {
    constexpr unsigned W = 8 * sizeof(T), L = 1024 / W, SZ = sizeof(T);
    __shared__ T       xs[1024];
    const uint64_t     v   = blockIdx.x;
    const unsigned     tid = threadIdx.x;
    for (unsigned k = tid; k < 1024; k += 128) xs[k] = x[v * 1024 + k];
    __syncthreads();
    const EncPage  pg = pages[vpage[v]];
    const uint32_t r  = (uint32_t)(v - pg.v0);
    uint8_t*       p  = out + pg.start;
    uint8_t*       ep = p + 8 + 4 * arr_c + 4ull * arr_c * (r ? r - 1 : 0);  // this vector's entry-point row
    if constexpr (MODE == 0) {
        uint8_t* dst = p + pg.off[0] + (uint64_t)r * 1024 * SZ;
        for (unsigned k = tid; k < 1024; k += 128) st_ua<T>(dst + k * SZ, xs[enc_orig(k)]);
        if (r && tid == 0) st_ua<uint32_t>(ep, r * 1024 * SZ);
    } else {
        const unsigned bw   = bw_in[v];
        const T        base = (T)base_in[v];
        constexpr unsigned o = MODE == 2 ? 1 : 0;  // bw buffer
        if (tid == 0) {
            p[pg.off[o] + r] = (uint8_t)bw;
            st_ua<T>(p + pg.off[o + 1] + (uint64_t)r * SZ, base);
            if (r) {
                if constexpr (MODE == 2) st_ua<uint32_t>(ep, r * 128), ep += 4;
                st_ua<uint32_t>(ep, r), st_ua<uint32_t>(ep + 4, r * SZ), st_ua<uint32_t>(ep + 8, (uint32_t)pk_rel[v]);
            }
        }
        if constexpr (MODE == 2)  // lane bases: the first L transposed values
            if (tid < L) st_ua<T>(p + pg.off[0] + (uint64_t)r * 128 + tid * SZ, xs[enc_orig(tid)]);
        if (tid < L && bw) {
            uint8_t*       pk   = p + pg.off[o + 2] + pk_rel[v];
            const uint64_t mask = bw >= 64 ? ~0ull : ((1ull << bw) - 1);
            using A             = std::conditional_t<sizeof(T) == 8, unsigned __int128, uint64_t>;
            A        acc = 0;
            unsigned nb = 0, w = 0;
            for (unsigned j = 0; j < W; ++j) {
                const uint64_t val = (uint64_t)(T)(enc_value<T, MODE>(xs, enc_orig(j * L + tid)) - base) & mask;
                acc |= (A)val << nb;
                nb += bw;
                while (nb >= W) {
                    st_ua<T>(pk + ((uint64_t)w * L + tid) * SZ, (T)acc);
                    acc >>= W, nb -= W, ++w;
                }
            }
        }
    }
}

// RLE (include/fls/cor/exp/rle/rle_exp.hpp:37-74, split of src/cor/lyt/vec.cpp:50-71): a vector's run values
// (one per change of value), and per value its u16 run index, transposed; VAR 1 stores the indexes, VAR 2 their
// 64 lane bases and the FFOR of their in-segment differences (16-value segments of u16), which are the run-start
// flags themselves: 0 at a segment start, else 1 where a run starts -- so base 0 and bw 0 or 1.
// Analysis, one wave per vector: runs[v] = number of runs, bw[v] = the difference FFOR's bit width (VAR 2).
template <typename T>
__global__ __launch_bounds__(256) void k_enc_rle_analyze(const T* __restrict__ x, uint64_t n_vec, uint8_t* __restrict__ bw,
                                                         uint64_t* __restrict__ runs) {
    const unsigned lane  = threadIdx.x % 64;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
    for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; v < n_vec; v += waves) {
        const T* xv  = x + v * 1024;
        unsigned n   = 0;
        bool     one = false;
        for (unsigned o = lane; o < 1024; o += 64) {
            const bool f = o == 0 || xv[o] != xv[o - 1];
            n += f;
            one |= f && (o % 16);
        }
        for (int d = 32; d >= 1; d >>= 1) n += __shfl_xor(n, d, 64);
        const bool any = __ballot(one) != 0;
        if (lane == 0) runs[v] = n, bw[v] = any ? 1 : 0;
    }
}

// One block of 128 threads per vector: run-start flags, run indexes by a block scan, then the page buffers.
template <typename T, int VAR>
__global__ __launch_bounds__(128) void k_enc_rle_pack(const T* __restrict__ x, const uint8_t* __restrict__ bw_in,
                                                      const uint64_t* __restrict__ rv_rel,
                                                      const uint64_t* __restrict__ pk_rel, const uint32_t* __restrict__ vpage,
                                                      const EncPage* __restrict__ pages, uint32_t arr_c,
                                                      uint8_t* __restrict__ out) {
    constexpr unsigned SZ = sizeof(T);
    __shared__ T        xs[1024];
    __shared__ uint16_t ps[1024];
    __shared__ unsigned part[2];
    const uint64_t      v   = blockIdx.x;
    const unsigned      tid = threadIdx.x, lane = tid % 64;
    for (unsigned k = tid; k < 1024; k += 128) xs[k] = x[v * 1024 + k];
    __syncthreads();
    // thread t owns values 8t .. 8t+7: local flags, then an exclusive scan over the block
    unsigned f8 = 0, c = 0;
    for (unsigned q = 0; q < 8; ++q) {
        const unsigned o = 8 * tid + q;
        const bool     f = o == 0 || xs[o] != xs[o - 1];
        f8 |= (unsigned)f << q;
        c += f;
    }
    unsigned inc = c;
    for (unsigned d = 1; d < 64; d <<= 1) {
        const unsigned y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
    }
    if (lane == 63) part[tid / 64] = inc;
    __syncthreads();
    unsigned run = inc - c + (tid >= 64 ? part[0] : 0);  // runs started before value 8t
    const EncPage  pg = pages[vpage[v]];
    const uint32_t r  = (uint32_t)(v - pg.v0);
    uint8_t*       p  = out + pg.start;
    uint8_t*       rv = p + pg.off[0] + rv_rel[v];
    for (unsigned q = 0; q < 8; ++q) {
        const unsigned o = 8 * tid + q;
        if ((f8 >> q) & 1u) st_ua<T>(rv + (uint64_t)run * SZ, xs[o]), ++run;
        ps[o] = (uint16_t)(run - 1);
    }
    __syncthreads();
    uint8_t* ep = p + 8 + 4 * arr_c + 4ull * arr_c * (r ? r - 1 : 0);
    if constexpr (VAR == 1) {
        for (unsigned k = tid; k < 1024; k += 128) st_ua<uint16_t>(p + pg.off[1] + (uint64_t)r * 2048 + 2 * k, ps[enc_orig(k)]);
        if (r && tid == 0) st_ua<uint32_t>(ep, (uint32_t)rv_rel[v]), st_ua<uint32_t>(ep + 4, r * 2048);
    } else {
        const unsigned bw = bw_in[v];
        if (tid < 64) st_ua<uint16_t>(p + pg.off[1] + (uint64_t)r * 128 + 2 * tid, ps[enc_orig(tid)]);
        if (tid == 0) {
            p[pg.off[2] + r] = (uint8_t)bw;
            st_ua<uint16_t>(p + pg.off[3] + (uint64_t)r * 2, 0);  // base: the minimum difference is 0
            if (r) {
                st_ua<uint32_t>(ep, (uint32_t)rv_rel[v]), st_ua<uint32_t>(ep + 4, r * 128);
                st_ua<uint32_t>(ep + 8, r), st_ua<uint32_t>(ep + 12, r * 2), st_ua<uint32_t>(ep + 16, (uint32_t)pk_rel[v]);
            }
        }
        if (bw && tid < 64) {  // bw 1: lane i's 16 differences (transposed j * 64 + i) form one u16 word
            unsigned word = 0;
            for (unsigned j = 0; j < 16; ++j) {
                const unsigned o = enc_orig(j * 64 + tid);
                word |= (unsigned)((o % 16) ? (ps[o] - ps[o - 1]) : 0) << j;
            }
            st_ua<uint16_t>(p + pg.off[4] + pk_rel[v] + 2 * tid, (uint16_t)word);
        }
    }
}

template <typename T>
hipError_t launch_enc(int mode, const void* x, uint64_t n_vec, uint8_t* bw, uint64_t* base, hipStream_t s, bool analyze,
                      const uint64_t* pk_rel, const uint32_t* vpage, const EncPage* pages, uint32_t arr_c,
                      uint8_t* out, bool sgn) {
    const T* xv = static_cast<const T*>(x);
    if (analyze) {
        const unsigned blocks = (unsigned)std::min<uint64_t>((n_vec + 3) / 4, 65536);
#define A(M, S) hipLaunchKernelGGL((k_enc_analyze<T, M, S>), dim3(blocks), dim3(256), 0, s, xv, n_vec, bw, base)
        if (mode == 1 && sgn) A(1, true);
        else if (mode == 1) A(1, false);
        else if (mode == 2 && sgn) A(2, true);
        else if (mode == 2) A(2, false);
#undef A
        else hipLaunchKernelGGL((k_enc_rle_analyze<T>), dim3(blocks), dim3(256), 0, s, xv, n_vec, bw, base);
    } else if (mode >= 3) {  // RLE: `base` carries the run-value offsets
#define R(V) hipLaunchKernelGGL((k_enc_rle_pack<T, V>), dim3((unsigned)n_vec), dim3(128), 0, s, xv, bw, base, pk_rel, vpage, pages, arr_c, out)
        if (mode == 3) R(1);
        else R(2);
#undef R
    } else {
#define P(M) hipLaunchKernelGGL((k_enc_pack<T, M>), dim3((unsigned)n_vec), dim3(128), 0, s, xv, bw, base, pk_rel, vpage, pages, arr_c, out)
        if (mode == 0) P(0);
        else if (mode == 1) P(1);
        else P(2);
#undef P
    }
    return hipGetLastError();
}

hipError_t launch_enc_w(unsigned W, int mode, const void* x, uint64_t n_vec, uint8_t* bw, uint64_t* base, hipStream_t s,
                        bool analyze, const uint64_t* pk_rel = nullptr, const uint32_t* vpage = nullptr,
                        const EncPage* pages = nullptr, uint32_t arr_c = 0, uint8_t* out = nullptr, bool sgn = false) {
    switch (W) {
    case 8: return launch_enc<uint8_t>(mode, x, n_vec, bw, base, s, analyze, pk_rel, vpage, pages, arr_c, out, sgn);
    case 16: return launch_enc<uint16_t>(mode, x, n_vec, bw, base, s, analyze, pk_rel, vpage, pages, arr_c, out, sgn);
    case 32: return launch_enc<uint32_t>(mode, x, n_vec, bw, base, s, analyze, pk_rel, vpage, pages, arr_c, out, sgn);
    default: return launch_enc<uint64_t>(mode, x, n_vec, bw, base, s, analyze, pk_rel, vpage, pages, arr_c, out, sgn);
    }
}

}  // namespace
}  // namespace fls

using namespace fls;

// The encoder's device scratch: per vector its bit width, FFOR base (RLE: run count), packed-buffer offset, page
// index and RLE run-value offset, plus the page table. `fixed` != nullptr: carved from that buffer (at most
// kEncFixedVec vectors and pages), else allocated on the stream and freed before returning.
namespace fls {
constexpr uint64_t kEncFixedVec = 16384;  // 2^24 values
constexpr size_t   kEncFixedBytes = kEncFixedVec * (1 + 8 + 8 + 4 + 8 + sizeof(EncPage)) + 6 * 256;
}  // namespace fls

namespace {
int encode_impl(const void* dev_values, uint64_t n, enum data_t type, uint8_t expr_id, uint8_t* dev_out,
                uint64_t capacity, uint64_t* out_size, void* stream, uint8_t* fixed) {
    if (!out_size || !dev_out || (n && !dev_values)) return fail(FLS_E_ARG, "null argument");
    Expr e;
    if (type == STR || !expr_from_id((uint8_t)type, expr_id, e)) return fail(FLS_E_UNSUPPORTED, "unknown expression");
    const int mode = e.exp_t == EXP_UNCOMPRESSED            ? 0
                     : (e.exp_t == EXP_FFOR && e.var == 1)  ? 1
                     : (e.exp_t == EXP_DELTA && e.var == 1) ? 2
                     : (e.exp_t == EXP_RLE && e.var == 1)   ? 3
                     : (e.exp_t == EXP_RLE && e.var == 2)   ? 4
                                                            : -1;
    if (mode < 0) return fail(FLS_E_UNSUPPORTED, "the GPU encoder writes UNCOMPRESSED, FFOR, DELTA and RLE pages (no patches, no chains)");
    // double RLE splits runs by double equality (+0.0 == -0.0, NaN != NaN): the host encoder writes those pages
    if (type == DOUBLE && mode) return fail(FLS_E_UNSUPPORTED, "the GPU encoder writes double columns UNCOMPRESSED only");
    if (n % 1024) return fail(FLS_E_UNSUPPORTED, "value count must be a multiple of 1024");
    const unsigned    W = dtype_bits((uint8_t)type), SZ = W / 8, arr_c = e.arr_c;
    const uint64_t    n_vec = n / 1024;
    hipStream_t       s     = (hipStream_t)stream;
    hipError_t        err;
    if (fixed && n_vec > kEncFixedVec)
        return fail(FLS_E_CAPACITY, "the allocation-free encoder takes at most 2^24 values per call");
    std::vector<uint8_t>  bw(n_vec, 0);
    std::vector<uint64_t> runs(mode >= 3 ? n_vec : 0);
    uint8_t*          d_bw   = nullptr;
    uint64_t*         d_base = nullptr;
    uint64_t*         d_pk   = nullptr;
    uint32_t*         d_vp   = nullptr;
    uint64_t*         d_rv   = nullptr;
    EncPage*          d_pg   = nullptr;
    auto              release = [&](int rc) {
        if (!fixed)
            for (void* q : {(void*)d_bw, (void*)d_base, (void*)d_pk, (void*)d_vp, (void*)d_pg, (void*)d_rv})
                if (q) (void)hipFreeAsync(q, s);
        (void)hipStreamSynchronize(s);
        return rc;
    };
    // scratch region k of `bytes` bytes: carved from the fixed buffer (regions of kEncFixedVec units, 256-aligned)
    // or allocated on the stream
    auto get = [&](auto** q, unsigned k, uint64_t unit, uint64_t count) -> hipError_t {
        if (!fixed) return hipMallocAsync((void**)q, unit * count, s);
        static const uint64_t kUnit[6] = {1, 8, 8, 4, 8, sizeof(EncPage)};
        uint64_t at = 0;
        for (unsigned r = 0; r < k; ++r) at += (kUnit[r] * kEncFixedVec + 255) / 256 * 256;
        *q = reinterpret_cast<std::remove_pointer_t<decltype(q)>>(fixed + at);
        return hipSuccess;
    };
    if (n_vec) {
        if ((err = get(&d_bw, 0, 1, n_vec)) != hipSuccess || (err = get(&d_base, 1, 8, n_vec)) != hipSuccess)
            return release(hip_fail(err, "hipMallocAsync"));
        if (mode) {
            if ((err = launch_enc_w(W, mode, dev_values, n_vec, d_bw, d_base, s, true, nullptr, nullptr, nullptr, 0,
                                    nullptr, type == INT64)) != hipSuccess ||
                (err = hipMemcpyAsync(bw.data(), d_bw, n_vec, hipMemcpyDeviceToHost, s)) != hipSuccess ||
                (mode >= 3 && (err = hipMemcpyAsync(runs.data(), d_base, 8 * n_vec, hipMemcpyDeviceToHost, s)) != hipSuccess) ||
                (err = hipStreamSynchronize(s)) != hipSuccess)
                return release(hip_fail(err, "analyze"));
        }
    }
    // page layout, as the host encoder's assembly loop (format.cpp encode_t step 2)
    auto vec_bytes = [&](uint64_t v, unsigned b) -> uint64_t {
        if (mode == 0) return 1024ull * SZ;
        if (mode >= 3) {  // RLE: [run values][positions] or [run values][lane bases][bw][base u16][packed]
            if (b == 0) return runs[v] * SZ;
            if (mode == 3) return 2048;
            return b == 1 ? 128 : b == 2 ? 1 : b == 3 ? 2 : 128ull * bw[v];
        }
        const unsigned o = mode == 2 ? 1 : 0;
        if (mode == 2 && b == 0) return 128;
        if (b == o) return 1;
        if (b == o + 1) return SZ;
        return 128ull * bw[v];
    };
    std::vector<EncPage>  pages;
    std::vector<uint8_t>  heads;  // per page: 8-byte header + arr_c offsets
    std::vector<uint32_t> vpage(n_vec);
    std::vector<uint64_t> pk_rel(n_vec), rv_rel(mode >= 3 ? n_vec : 0);
    const uint64_t        page_c = n_vec ? (n_vec + kMaxVecPerPage - 1) / kMaxVecPerPage : 1;
    const uint64_t        per_pg = n_vec ? (n_vec + page_c - 1) / page_c : 0;
    uint64_t              v = 0, pos = 0;
    for (;;) {
        const uint64_t v0 = v, stop = std::min(n_vec, v + per_pg);
        uint64_t       bufs[6] = {0, 0, 0, 0, 0, 0}, eps = 0, vec_c = 0;
        for (; v < stop; ++v) {
            uint64_t add = 4ull * arr_c, cur = 8 + 4ull * arr_c + eps;
            for (unsigned b = 0; b < arr_c; ++b) add += vec_bytes(v, b), cur += bufs[b];
            if (vec_c && cur + add >= kDataPageMax) break;  // TryFlushVec limit (compressor.cpp:47)
            if (vec_c) eps += 4ull * arr_c;
            pk_rel[v] = mode == 0 ? 0 : bufs[arr_c - 1];
            if (mode >= 3) rv_rel[v] = bufs[0];
            vpage[v]  = (uint32_t)pages.size();
            for (unsigned b = 0; b < arr_c; ++b) bufs[b] += vec_bytes(v, b);
            ++vec_c;
        }
        const bool more = v < n_vec;
        EncPage    pg {};
        pg.start = pos, pg.v0 = v0;
        uint64_t off = 8 + 4ull * arr_c + eps;
        const uint8_t hdr[8] = {e.exp_t, e.var, 0, 0, (uint8_t)(vec_c & 0xff), (uint8_t)(vec_c >> 8), (uint8_t)more, 0};
        heads.insert(heads.end(), hdr, hdr + 8);
        for (unsigned b = 0; b < arr_c; ++b) {
            pg.off[b]      = (uint32_t)off;
            const uint32_t o32 = (uint32_t)off;
            heads.insert(heads.end(), (const uint8_t*)&o32, (const uint8_t*)&o32 + 4);
            off += bufs[b];
        }
        pages.push_back(pg);
        pos += off;
        if (!more) break;
    }
    *out_size = pos;
    if (pos > capacity) return release(fail(FLS_E_CAPACITY, "encoded column exceeds capacity"));
    if (fixed && pages.size() > kEncFixedVec) return release(fail(FLS_E_CAPACITY, "too many pages"));
    // headers and buffer offsets (host -> device), then every vector's buffers
    const uint64_t hb = 8 + 4ull * arr_c;
    for (size_t q = 0; q < pages.size(); ++q)
        if ((err = hipMemcpyAsync(dev_out + pages[q].start, heads.data() + q * hb, hb, hipMemcpyHostToDevice, s)) != hipSuccess)
            return release(hip_fail(err, "hipMemcpyAsync"));
    if (n_vec) {
        if ((err = get(&d_pk, 2, 8, n_vec)) != hipSuccess || (err = get(&d_vp, 3, 4, n_vec)) != hipSuccess ||
            (err = get(&d_pg, 5, sizeof(EncPage), pages.size())) != hipSuccess)
            return release(hip_fail(err, "hipMallocAsync"));
        if ((err = hipMemcpyAsync(d_pk, pk_rel.data(), 8 * n_vec, hipMemcpyHostToDevice, s)) != hipSuccess ||
            (err = hipMemcpyAsync(d_vp, vpage.data(), 4 * n_vec, hipMemcpyHostToDevice, s)) != hipSuccess ||
            (err = hipMemcpyAsync(d_pg, pages.data(), sizeof(EncPage) * pages.size(), hipMemcpyHostToDevice, s)) != hipSuccess)
            return release(hip_fail(err, "hipMemcpyAsync"));
        if (mode >= 3 && ((err = get(&d_rv, 4, 8, n_vec)) != hipSuccess ||
                          (err = hipMemcpyAsync(d_rv, rv_rel.data(), 8 * n_vec, hipMemcpyHostToDevice, s)) != hipSuccess))
            return release(hip_fail(err, "run-value offsets"));
        if ((err = launch_enc_w(W, mode, dev_values, n_vec, d_bw, mode >= 3 ? d_rv : d_base, s, false, d_pk, d_vp, d_pg,
                                arr_c, dev_out)) != hipSuccess)
            return release(hip_fail(err, "kernel launch"));
    }
    return release(FLS_OK);
}
}  // namespace

// Device scratch of fls_gpu_encode_noalloc: one copy per GPU, held until the process ends. Host code never uses
// this name as an address: each device's address comes from hipGetSymbolAddress (scratch_on_device).
static __device__ __align__(256) uint8_t scratch[fls::kEncFixedBytes];

namespace {
// The address of this device's copy of `scratch`, looked up once per device with that device current.
int scratch_on_device(uint8_t** out) {
    static std::mutex            mu;
    static std::vector<uint8_t*> by_dev;
    int                          dev = 0;
    hipError_t                   e   = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)dev >= by_dev.size()) by_dev.resize(dev + 1, nullptr);
    if (!by_dev[dev]) {
        void* p = nullptr;
        if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return hip_fail(e, "hipGetSymbolAddress");
        by_dev[dev] = static_cast<uint8_t*>(p);
    }
    *out = by_dev[dev];
    return FLS_OK;
}
}  // namespace

extern "C" int fls_gpu_encode(const void* dev_values, uint64_t n, enum data_t type, uint8_t expr_id, uint8_t* dev_out,
                              uint64_t capacity, uint64_t* out_size, void* stream) {
    return encode_impl(dev_values, n, type, expr_id, dev_out, capacity, out_size, stream, nullptr);
}

extern "C" int fls_gpu_encode_noalloc(const void* dev_values, uint64_t n, enum data_t type, uint8_t expr_id,
                                      uint8_t* dev_out, uint64_t capacity, uint64_t* out_size, void* stream) {
    if (n / 1024 > kEncFixedVec) return fail(FLS_E_CAPACITY, "the allocation-free encoder takes at most 2^24 values per call");
    uint8_t* base = nullptr;
    if (int rc = scratch_on_device(&base)) return rc;
    return encode_impl(dev_values, n, type, expr_id, dev_out, capacity, out_size, stream, base);
}
