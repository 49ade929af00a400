// k_fsst.hip -- FSST string pages on gfx950.
//
// Page per vector (include/fls/cor/exp/fsst/fsst_exp.hpp:35-52, src/cor/lyt/vec.cpp:160-207): buffer 0 holds the
// 1024 compressed strings back to back, buffer 1 their cumulative end offsets (u32) in transposed order.
// Decoding follows include/fls/encoder/assert_eq.hpp:166-210: untranspose the offsets, then decode each string
// with fsst_decompress (include/fls/cor/prm/fsst/fsst.h:152-258): code < 255 emits symbol[code] (len[code] bytes),
// code 255 emits the next byte literally.
//
// One 256-thread block per vector, 4 strings per thread; symbol table in LDS.
#include "../../include/fls_gpu.h"
#include "gpu_internal.hpp"
#include "page_kernels.hpp"

namespace fls {

namespace {

constexpr unsigned kBlock        = 256;
constexpr unsigned kStrPerThread = 1024 / kBlock;

// transposed index of original position i (inverse of untranspose_i): i = r + 8*ORDER[b] + 64c -> 128r + 16b + c
__device__ __forceinline__ unsigned transposed_of(unsigned i) {
    const unsigned c = i >> 6, b = order8((i >> 3) & 7), r = i & 7;
    return 128 * r + 16 * b + c;
}

__device__ __forceinline__ void string_range(const uint8_t* ends, unsigned i, uint32_t& s, uint32_t& e) {
    const u32_ua* E = reinterpret_cast<const u32_ua*>(ends);
    e               = E[transposed_of(i)];
    s               = i ? E[transposed_of(i - 1)] : 0u;
    if (s > e) s = e;  // malformed offsets decode to an empty string
}

__global__ __launch_bounds__(kBlock) void k_fsst_len(const uint8_t* __restrict__ page, uint32_t vec_c,
                                                     const FsstTableDev* __restrict__ table, uint64_t first_string,
                                                     uint64_t* __restrict__ offsets, uint64_t* __restrict__ vec_tot,
                                                     uint64_t first_vec) {
    __shared__ uint8_t  len[256];
    __shared__ uint64_t part[kBlock];
    const unsigned      v = blockIdx.x, t = threadIdx.x;
    if (t < 256) len[t] = table->len[t];
    __syncthreads();
    const uint8_t* cb   = vbuf(page, 2, v, 0);
    const uint8_t* ends = vbuf(page, 2, v, 1);
    uint64_t       l[kStrPerThread], run = 0;
    for (unsigned k = 0; k < kStrPerThread; ++k) {
        const unsigned i = t * kStrPerThread + k;
        uint32_t       s, e;
        string_range(ends, i, s, e);
        uint64_t n = 0;
        for (uint32_t q = s; q < e;) {
            const unsigned c = cb[q];
            if (c == 255) n += 1, q += 2;
            else n += len[c], q += 1;
        }
        run += n;
        l[k] = run;
    }
    // block exclusive scan of per-thread totals
    part[t] = run;
    __syncthreads();
    for (unsigned d = 1; d < kBlock; d <<= 1) {
        const uint64_t x = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    const uint64_t before = t ? part[t - 1] : 0;
    uint64_t*      o      = offsets + first_string + (uint64_t)v * 1024 + 1;
    for (unsigned k = 0; k < kStrPerThread; ++k) o[t * kStrPerThread + k] = before + l[k];
    if (t == kBlock - 1) vec_tot[first_vec + v] = part[kBlock - 1];
}

// Single-block exclusive scan over the per-vector totals (n <= a few million).
__global__ __launch_bounds__(1024) void k_scan_vec(uint64_t* __restrict__ tot, uint64_t n, uint64_t* __restrict__ offsets) {
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

__global__ __launch_bounds__(kBlock) void k_fsst_fix(uint64_t* __restrict__ offsets, const uint64_t* __restrict__ vbase,
                                                     uint64_t n_vec) {
    const uint64_t v = blockIdx.x;
    if (v >= n_vec) return;
    const uint64_t b = vbase[v];
    for (unsigned k = threadIdx.x; k < 1024; k += kBlock) offsets[v * 1024 + 1 + k] += b;
}

__global__ __launch_bounds__(kBlock) void k_fsst_decode(const uint8_t* __restrict__ page, uint32_t vec_c,
                                                        const FsstTableDev* __restrict__ table, uint64_t first_string,
                                                        const uint64_t* __restrict__ offsets, uint8_t* __restrict__ out,
                                                        uint64_t cap, uint32_t* err) {
    __shared__ uint64_t sym[256];
    __shared__ uint8_t  len[256];
    const unsigned      v = blockIdx.x, t = threadIdx.x;
    sym[t] = table->symbol[t];
    len[t] = table->len[t];
    __syncthreads();
    const uint8_t*  cb   = vbuf(page, 2, v, 0);
    const uint8_t*  ends = vbuf(page, 2, v, 1);
    const uint64_t* o    = offsets + first_string + (uint64_t)v * 1024;
    for (unsigned k = 0; k < kStrPerThread; ++k) {
        const unsigned i = t * kStrPerThread + k;
        uint32_t       s, e;
        string_range(ends, i, s, e);
        uint64_t       w = o[i];
        const uint64_t stop = o[i + 1];
        if (stop > cap) {
            atomicOr(err, 4u);
            continue;
        }
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
    if ((e = hipMalloc(&sp.d_vec_tot, (sp.n_vec + 1) * sizeof(uint64_t))) != hipSuccess) return hip_fail(e, "hipMalloc");
    return FLS_OK;
}

void strings_release(StringPlan& sp) {
    if (sp.d_tables) (void)hipFree(sp.d_tables);
    if (sp.d_vec_tot) (void)hipFree(sp.d_vec_tot);
    sp = StringPlan {};
}

int strings_offsets(const Column& col, StringPlan& sp, const uint8_t* dev_enc, uint64_t* dev_offsets, hipStream_t s) {
    for (size_t i = 0; i < col.pages.size(); ++i) {
        const PageEntry& pe = col.pages[i];
        if (!pe.vec_c) continue;
        hipLaunchKernelGGL(k_fsst_len, dim3(pe.vec_c), dim3(kBlock), 0, s, dev_enc + pe.page_off, pe.vec_c,
                           sp.d_tables + i, pe.out_index, dev_offsets, sp.d_vec_tot, pe.out_index / 1024);
    }
    if (sp.n_vec) {
        hipLaunchKernelGGL(k_scan_vec, dim3(1), dim3(1024), 0, s, sp.d_vec_tot, sp.n_vec, dev_offsets);
        hipLaunchKernelGGL(k_fsst_fix, dim3((unsigned)sp.n_vec), dim3(kBlock), 0, s, dev_offsets, sp.d_vec_tot, sp.n_vec);
    } else {
        hipError_t e = hipMemsetAsync(dev_offsets, 0, 8, s);
        if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "kernel launch");
}

int strings_decode(const Column& col, StringPlan& sp, const uint8_t* dev_enc, const uint64_t* dev_offsets,
                   uint8_t* dev_bytes, uint64_t cap, hipStream_t s) {
    static uint32_t* d_err = nullptr;  // error word for string pages (never read back on the hot path)
    if (!d_err) {
        hipError_t e = hipMalloc(&d_err, 4);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc");
        (void)hipMemset(d_err, 0, 4);
    }
    for (size_t i = 0; i < col.pages.size(); ++i) {
        const PageEntry& pe = col.pages[i];
        if (!pe.vec_c) continue;
        hipLaunchKernelGGL(k_fsst_decode, dim3(pe.vec_c), dim3(kBlock), 0, s, dev_enc + pe.page_off, pe.vec_c,
                           sp.d_tables + i, pe.out_index, dev_offsets, dev_bytes, cap, d_err);
    }
    hipError_t e = hipGetLastError();
    (void)sp;
    return e == hipSuccess ? FLS_OK : hip_fail(e, "kernel launch");
}

}  // namespace fls
