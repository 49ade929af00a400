// k_fsst.hip -- FSST string pages on gfx950.
//
// Page per vector (include/fls/cor/exp/fsst/fsst_exp.hpp:35-52, src/cor/lyt/vec.cpp:160-207): buffer 0 holds the
// 1024 compressed strings back to back, buffer 1 their cumulative end offsets (u32) in transposed order.
// Decoding follows include/fls/encoder/assert_eq.hpp:166-210: untranspose the offsets, then decode each string
// with fsst_decompress (include/fls/cor/prm/fsst/fsst.h:152-258): code < 255 emits symbol[code] (len[code] bytes),
// code 255 emits the next byte literally.
//
// A block of 512 threads owns one half-vector segment (512 strings, one per thread). The segment's compressed
// bytes are staged in LDS with a coalesced copy, strings are decoded from LDS into an LDS output buffer, and the
// decoded segment is written out with consecutive lanes storing consecutive bytes. Segments larger than the LDS
// budget decode straight from / to global memory.
#include "../../include/fls_gpu.h"
#include "gpu_internal.hpp"
#include "page_kernels.hpp"

namespace fls {

namespace {

constexpr unsigned kSeg      = 512;        // strings per segment (= threads per block)
constexpr unsigned kSegPerV  = 1024 / kSeg;
constexpr uint32_t kCompLds  = 20 * 1024;  // staged compressed bytes per segment
constexpr uint32_t kOutLds   = 24 * 1024;  // staged decoded bytes per segment

// transposed index of original position i (inverse of untranspose_i): i = r + 8*ORDER[b] + 64c -> 128r + 16b + c
__device__ __forceinline__ unsigned transposed_of(unsigned i) {
    const unsigned c = i >> 6, b = order8((i >> 3) & 7), r = i & 7;
    return 128 * r + 16 * b + c;
}

__device__ __forceinline__ uint32_t end_of(const uint8_t* ends, int i) {  // end offset of string i, 0 for i = -1
    return i < 0 ? 0u : reinterpret_cast<const u32_ua*>(ends)[transposed_of((unsigned)i)];
}

// Exclusive scan of one u64 per thread over a block of kSeg threads; returns the block total in *total.
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t x, uint64_t* total) {
    __shared__ uint64_t wsum[kSeg / 64];
    const unsigned      lane = threadIdx.x % 64, w = threadIdx.x / 64;
    uint64_t            inc  = x;
    for (unsigned d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint64_t before = 0, all = 0;
    for (unsigned k = 0; k < kSeg / 64; ++k) {
        const uint64_t s = wsum[k];
        if (k < w) before += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + inc - x;
}

// Stages compressed bytes [s0, s1) of a segment into LDS (16-byte loads from the unaligned source, aligned LDS
// stores, the last n % 16 bytes one by one so nothing past s1 is read); returns whether they fit.
__device__ __forceinline__ bool stage_bytes(const uint8_t* cb, uint32_t s0, uint32_t s1, uint8_t* lds) {
    const uint32_t n = s1 - s0;
    if (n > kCompLds) return false;
    const uint32_t n16 = n & ~15u;
    for (uint32_t k = 16 * threadIdx.x; k < n16; k += 16 * kSeg)
        *reinterpret_cast<v4u*>(lds + k) = *reinterpret_cast<const v4u_ua*>(cb + s0 + k);
    for (uint32_t k = n16 + threadIdx.x; k < n; k += kSeg) lds[k] = cb[s0 + k];
    return true;
}

__global__ __launch_bounds__(kSeg) void k_fsst_len(const uint8_t* __restrict__ page, const FsstTableDev* __restrict__ table,
                                                   uint64_t first_string, uint64_t* __restrict__ offsets,
                                                   uint64_t* __restrict__ seg_tot, uint64_t first_seg) {
    __shared__ uint8_t len[256];
    __shared__ __attribute__((aligned(16))) uint8_t cbuf[kCompLds];
    const unsigned     v = blockIdx.x / kSegPerV, h = blockIdx.x % kSegPerV, i = h * kSeg + threadIdx.x;
    if (threadIdx.x < 256) len[threadIdx.x] = table->len[threadIdx.x];
    const uint8_t* cb   = vbuf(page, 2, v, 0);
    const uint8_t* ends = vbuf(page, 2, v, 1);
    const uint32_t s0 = end_of(ends, (int)(h * kSeg) - 1), s1 = end_of(ends, (int)(h * kSeg + kSeg) - 1);
    const bool     in_lds = stage_bytes(cb, s0, s1 > s0 ? s1 : s0, cbuf);
    __syncthreads();
    uint32_t s = end_of(ends, (int)i - 1), e = end_of(ends, (int)i);
    if (s > e) s = e;  // malformed offsets decode to an empty string
    uint64_t n = 0;
    if (in_lds && s >= s0 && e <= s1) {
        for (uint32_t q = s - s0; q < e - s0;) {
            const unsigned c   = cbuf[q];
            const bool     esc = c == 255;
            n += esc ? 1u : len[c];
            q += esc ? 2u : 1u;
        }
    } else {
        for (uint32_t q = s; q < e;) {
            const unsigned c = cb[q];
            if (c == 255) n += 1, q += 2;
            else n += len[c], q += 1;
        }
    }
    uint64_t total;
    const uint64_t before = block_exclusive_scan(n, &total);
    offsets[first_string + (uint64_t)v * 1024 + i + 1] = before + n;  // relative to the segment; fixed below
    if (threadIdx.x == 0) seg_tot[first_seg + blockIdx.x] = total;
}

// Single-block exclusive scan over the per-segment totals.
__global__ __launch_bounds__(1024) void k_scan_seg(uint64_t* __restrict__ tot, uint64_t n, uint64_t* __restrict__ offsets) {
    __shared__ uint64_t part[1024];
    const unsigned      t     = threadIdx.x;
    const uint64_t      chunk = (n + 1023) / 1024, a = t * chunk, b = a + chunk < n ? a + chunk : n;
    uint64_t            s     = 0;
    for (uint64_t k = a; k < b; ++k) s += tot[k];
    part[t] = s;
    __syncthreads();
    for (unsigned d = 1; d < 1024; d <<= 1) {
        const uint64_t x = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint64_t run = t ? part[t - 1] : 0;
    for (uint64_t k = a; k < b; ++k) {
        const uint64_t x = tot[k];
        tot[k]           = run;
        run += x;
    }
    if (t == 0) offsets[0] = 0;
}

__global__ __launch_bounds__(kSeg) void k_fsst_fix(uint64_t* __restrict__ offsets, const uint64_t* __restrict__ sbase,
                                                   uint64_t n_seg) {
    const uint64_t g = blockIdx.x;
    if (g < n_seg) offsets[g * kSeg + 1 + threadIdx.x] += sbase[g];
}

// Writes a string's decoded bytes to [w, wend) of an LDS buffer a word at a time: bytes gather in a 128-bit
// pending register; single bytes are written until w is 4-byte aligned, then whole 32-bit words. Strings own
// disjoint byte ranges, so no write leaves its own string.
struct WordWriter {
    uint8_t*          buf;
    uint32_t          w, wend;
    unsigned __int128 pend = 0;
    unsigned          pb   = 0;  // pending bytes
    __device__ __forceinline__ void flush() {
        while (pb && (w & 3) && w < wend) {
            buf[w++] = (uint8_t)pend;
            pend >>= 8, --pb;
        }
        while (pb >= 4 && w + 4 <= wend) {
            *reinterpret_cast<uint32_t*>(buf + w) = (uint32_t)pend;
            w += 4, pend >>= 32, pb -= 4;
        }
    }
    __device__ __forceinline__ void put(uint64_t x, unsigned n) {  // n <= 8 bytes of x
        pend |= (unsigned __int128)x << (8 * pb);
        pb += n;
        flush();
    }
    __device__ __forceinline__ void finish() {
        while (pb && w < wend) {
            buf[w++] = (uint8_t)pend;
            pend >>= 8, --pb;
        }
    }
};

// Decodes a staged segment (n compressed bytes in cbuf, this thread's string starting at my_start) into obuf,
// splitting the bytes evenly over the block's threads. A byte is a literal when it follows an odd number of 255
// bytes inside its string (a run of escapes starts with a code, so codes and literals alternate along it); a code
// c < 255 emits symbol[c], an escape emits the byte after it (0 when the string ends first), a literal emits
// nothing -- the same lengths k_fsst_len counts per string. Each thread's output range comes from a block scan of
// its byte share; it assembles its bytes in a 64-bit register and stores whole aligned 32-bit words, OR-ing into
// the (pre-zeroed) first and last words it shares with its neighbours. Returns false (nothing written) when the
// total disagrees with the segment's offsets; every thread returns the same value.
__device__ __forceinline__ bool decode_segment_bytes(const uint64_t* sym, const uint8_t* len, const uint8_t* cbuf,
                                                     uint32_t n, uint32_t my_start, uint8_t* obuf, uint64_t n_out) {
    __shared__ uint32_t sbits[kCompLds / 32 + 1];  // string starts over positions [0, n]
    const unsigned      tid = threadIdx.x;
    for (uint32_t k = tid; k <= n / 32; k += kSeg) sbits[k] = 0u;
    __syncthreads();
    atomicOr(&sbits[my_start >> 5], 1u << (my_start & 31));
    if (tid == 0) atomicOr(&sbits[n >> 5], 1u << (n & 31));
    __syncthreads();
    auto is_start = [&](uint32_t p) { return ((sbits[p >> 5] >> (p & 31)) & 1u) != 0; };
    const uint32_t C = (n + kSeg - 1) / kSeg, q0 = tid * C < n ? tid * C : n, q1 = q0 + C < n ? q0 + C : n;
    bool lit0 = false;
    if (q0 < q1 && !is_start(q0)) {
        uint32_t p = q0;
        while (!is_start(p) && cbuf[p - 1] == 255) --p;
        lit0 = ((q0 - p) & 1u) != 0;
    }
    uint32_t tot = 0;
    bool     lit = lit0;
    for (uint32_t q = q0; q < q1; ++q) {
        const unsigned c = cbuf[q];
        if (lit) {
            lit = false;
        } else if (c == 255) {
            tot += 1;
            lit = !is_start(q + 1);
        } else {
            tot += len[c];
        }
    }
    uint64_t total;
    const uint32_t base = (uint32_t)block_exclusive_scan(tot, &total);
    if (total != n_out) return false;
    uint32_t* words = reinterpret_cast<uint32_t*>(obuf);
    const uint32_t first = base >> 2, last = (base + tot - 1) >> 2;
    if (tot) words[first] = 0u, words[last] = 0u;
    __syncthreads();
    if (tot) {
        uint32_t w   = first;
        uint64_t acc = 0;
        unsigned ab  = base & 3u;  // pending bytes in acc (the leading ones are the neighbour's: zeros)
        auto put_word = [&](uint32_t x) {
            if (w == first || w == last) atomicOr(&words[w], x);
            else words[w] = x;
            ++w;
        };
        lit = lit0;
        for (uint32_t q = q0; q < q1; ++q) {
            const unsigned c = cbuf[q];
            uint64_t       x;
            unsigned       k;
            if (lit) {
                lit = false;
                continue;
            } else if (c == 255) {
                lit = !is_start(q + 1);
                x = lit ? cbuf[q + 1] : 0u, k = 1;
            } else {
                x = sym[c], k = len[c];
            }
            const uint64_t lo  = acc | (x << (8 * ab));
            const uint64_t hi  = ab ? x >> (64 - 8 * ab) : 0ull;
            const unsigned all = ab + k;
            if (all >= 8) {
                put_word((uint32_t)lo), put_word((uint32_t)(lo >> 32));
                acc = hi, ab = all - 8;
            } else if (all >= 4) {
                put_word((uint32_t)lo);
                acc = lo >> 32, ab = all - 4;
            } else {
                acc = lo, ab = all;
            }
        }
        if (ab) atomicOr(&words[w], (uint32_t)acc);  // the last, partial word
    }
    return true;
}

__global__ __launch_bounds__(kSeg) void k_fsst_decode(const uint8_t* __restrict__ page, const FsstTableDev* __restrict__ table,
                                                      uint64_t first_string, const uint64_t* __restrict__ offsets,
                                                      uint8_t* __restrict__ out, uint64_t cap, uint32_t* err) {
    __shared__ uint64_t sym[256];
    __shared__ uint8_t  len[256];
    __shared__ __attribute__((aligned(16))) uint8_t cbuf[kCompLds];
    __shared__ __attribute__((aligned(16))) uint8_t obuf[kOutLds];
    const unsigned      v = blockIdx.x / kSegPerV, h = blockIdx.x % kSegPerV, i = h * kSeg + threadIdx.x;
    if (threadIdx.x < 256) sym[threadIdx.x] = table->symbol[threadIdx.x], len[threadIdx.x] = table->len[threadIdx.x];
    const uint8_t*  cb   = vbuf(page, 2, v, 0);
    const uint8_t*  ends = vbuf(page, 2, v, 1);
    const uint64_t* o    = offsets + first_string + (uint64_t)v * 1024;
    const uint32_t  s0 = end_of(ends, (int)(h * kSeg) - 1), s1 = end_of(ends, (int)(h * kSeg + kSeg) - 1);
    const uint64_t  o0 = o[h * kSeg], o1 = o[h * kSeg + kSeg];
    const bool      in_lds  = stage_bytes(cb, s0, s1 > s0 ? s1 : s0, cbuf);
    const bool      out_lds = in_lds && o1 >= o0 && o1 - o0 <= kOutLds && o1 <= cap;
    uint32_t s = end_of(ends, (int)i - 1), e = end_of(ends, (int)i);
    const uint64_t w0 = o[i], stop = o[i + 1];
    // Byte-parallel form: when the segment's strings are well formed, every thread decodes an equal share of the
    // segment's compressed bytes instead of one whole string (a wave otherwise runs as long as its longest string).
    const bool fast = __syncthreads_and(out_lds && s0 <= s && s <= e && e <= s1 && w0 <= stop && stop <= cap) != 0;
    if (fast && decode_segment_bytes(sym, len, cbuf, s1 - s0, s - s0, obuf, o1 - o0)) {
        // decoded into obuf; the copy-out follows
    } else {
    if (s > e) s = e;
    if (stop > cap || stop < w0) {
        atomicOr(err, 4u);
    } else if (out_lds && s >= s0 && e <= s1) {
        WordWriter ww {obuf, (uint32_t)(w0 - o0), (uint32_t)(stop - o0)};
        for (uint32_t q = s - s0; q < e - s0 && ww.w < ww.wend;) {
            const unsigned c = cbuf[q];
            if (c == 255) {
                ww.put(q + 1 < e - s0 ? cbuf[q + 1] : 0, 1);
                q += 2;
            } else {
                ww.put(sym[c], len[c]);
                q += 1;
            }
        }
        ww.finish();
    } else {
        uint64_t w = w0;
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
    }
    if (out_lds) {
        __syncthreads();
        // copy-out: 16 bytes per thread per step (unaligned global stores), then the tail byte by byte
        const uint32_t n = (uint32_t)(o1 - o0), n16 = n & ~15u;
        uint8_t*       dst = out + o0;
        for (uint32_t k = 16 * threadIdx.x; k < n16; k += 16 * kSeg)
            *reinterpret_cast<v4u_ua*>(dst + k) = *reinterpret_cast<const v4u*>(obuf + k);
        for (uint32_t k = n16 + threadIdx.x; k < n; k += kSeg) dst[k] = obuf[k];
    }
}

}  // namespace

int strings_prepare(const Column& col, const uint8_t* host, StringPlan& sp) {
    (void)host;
    std::vector<FsstTableDev> tabs(col.pages.size());
    for (size_t i = 0; i < col.pages.size(); ++i) {
        std::memset(&tabs[i], 0, sizeof(FsstTableDev));
        for (int c = 0; c < 255; ++c) tabs[i].symbol[c] = col.pages[i].fsst.symbol[c], tabs[i].len[c] = col.pages[i].fsst.len[c];
        tabs[i].len[255] = 1;
    }
    sp.n_vec = col.n_values / 1024;
    hipError_t e;
    if ((e = hipMalloc(&sp.d_tables, tabs.size() * sizeof(FsstTableDev) + 1)) != hipSuccess) return hip_fail(e, "hipMalloc");
    if ((e = hipMemcpy(sp.d_tables, tabs.data(), tabs.size() * sizeof(FsstTableDev), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "hipMemcpy");
    if ((e = hipMalloc(&sp.d_vec_tot, (sp.n_vec * kSegPerV + 1) * sizeof(uint64_t))) != hipSuccess)
        return hip_fail(e, "hipMalloc");
    if ((e = hipMalloc(&sp.d_err, 4)) != hipSuccess) return hip_fail(e, "hipMalloc");
    if ((e = hipMemset(sp.d_err, 0, 4)) != hipSuccess) return hip_fail(e, "hipMemset");
    return FLS_OK;
}

void strings_release(StringPlan& sp) {
    if (sp.d_tables) (void)hipFree(sp.d_tables);
    if (sp.d_vec_tot) (void)hipFree(sp.d_vec_tot);
    if (sp.d_err) (void)hipFree(sp.d_err);
    sp = StringPlan {};
}

int strings_offsets(const Column& col, StringPlan& sp, const uint8_t* dev_enc, uint64_t* dev_offsets, hipStream_t s) {
    for (size_t i = 0; i < col.pages.size(); ++i) {
        const PageEntry& pe = col.pages[i];
        if (!pe.vec_c) continue;
        hipLaunchKernelGGL(k_fsst_len, dim3(pe.vec_c * kSegPerV), dim3(kSeg), 0, s, dev_enc + pe.page_off, sp.d_tables + i,
                           pe.out_index, dev_offsets, sp.d_vec_tot, (pe.out_index / 1024) * kSegPerV);
    }
    const uint64_t n_seg = sp.n_vec * kSegPerV;
    if (n_seg) {
        hipLaunchKernelGGL(k_scan_seg, dim3(1), dim3(1024), 0, s, sp.d_vec_tot, n_seg, dev_offsets);
        hipLaunchKernelGGL(k_fsst_fix, dim3((unsigned)n_seg), dim3(kSeg), 0, s, dev_offsets, sp.d_vec_tot, n_seg);
    } else {
        hipError_t e = hipMemsetAsync(dev_offsets, 0, 8, s);
        if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "kernel launch");
}

int strings_decode(const Column& col, StringPlan& sp, const uint8_t* dev_enc, const uint64_t* dev_offsets,
                   uint8_t* dev_bytes, uint64_t cap, hipStream_t s) {
    for (size_t i = 0; i < col.pages.size(); ++i) {
        const PageEntry& pe = col.pages[i];
        if (!pe.vec_c) continue;
        hipLaunchKernelGGL(k_fsst_decode, dim3(pe.vec_c * kSegPerV), dim3(kSeg), 0, s, dev_enc + pe.page_off,
                           sp.d_tables + i, pe.out_index, dev_offsets, dev_bytes, cap, sp.d_err);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "kernel launch");
}

int strings_check(StringPlan& sp, hipStream_t s) {
    uint32_t   h = 0;
    hipError_t e = hipMemcpyAsync(&h, sp.d_err, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "check");
    return h ? fail(FLS_E_DATA, "string offsets exceed the output capacity") : FLS_OK;
}

}  // namespace fls
