// select.hip -- selection vectors of the predicate pushdown: bitmaps to index arrays, and their conjunction.
//
// The reference keeps a selection per 1024-value vector as SelectionDS (include/fls/expression/selection_ds.hpp):
// an index array idx_arr with n_selected_values, or a 16-word bitmap. The predicates (fls_gpu_plan_select) write
// bitmaps; these kernels give the index-array form (the ascending positions of the set bits) and the AND of two
// selections (and_selection_ds_func, src/primitive/predicate/equal.cpp:96-118, a merge of two sorted index arrays,
// is a word-wise AND of bitmaps).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fls_gpu.h"
#include "gpu_internal.hpp"

namespace fls {
namespace {

// One wave per vector: lane j < 16 owns word j; a shuffle scan of the word popcounts gives each word's first
// output slot, then the lane writes the positions of its set bits.
__global__ __launch_bounds__(256) void k_sel_indices(const uint64_t* __restrict__ bitmap, uint64_t n_vec,
                                                     uint16_t* __restrict__ idx, uint32_t* __restrict__ counts) {
    const unsigned lane = threadIdx.x % 64;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
    for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; v < n_vec; v += waves) {
        uint64_t       w    = lane < 16 ? bitmap[v * 16 + lane] : 0ull;
        const unsigned own  = (unsigned)__popcll(w);
        unsigned       incl = own;
        for (unsigned d = 1; d < 16; d <<= 1) {
            const unsigned y = __shfl_up(incl, d, 64);
            if (lane >= d) incl += y;
        }
        unsigned  k   = incl - own;
        uint16_t* out = idx + v * 1024;
        while (w) {
            out[k++] = (uint16_t)(64 * lane + (unsigned)__builtin_ctzll(w));
            w &= w - 1;
        }
        if (lane == 15 && counts) counts[v] = incl;
    }
}

__global__ __launch_bounds__(256) void k_sel_and(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                                                 uint64_t* __restrict__ out, uint64_t n_vec, uint32_t* __restrict__ counts) {
    const unsigned lane = threadIdx.x % 64;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
    for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; v < n_vec; v += waves) {
        const uint64_t w = lane < 16 ? a[v * 16 + lane] & b[v * 16 + lane] : 0ull;
        if (lane < 16) out[v * 16 + lane] = w;
        unsigned c = (unsigned)__popcll(w);
        for (unsigned d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
        if (lane == 0 && counts) counts[v] = c;
    }
}

unsigned waves_grid(uint64_t n_vec) {
    const uint64_t blocks = (n_vec + 3) / 4;  // four waves (vectors) per block
    return (unsigned)(blocks < 65536 ? (blocks ? blocks : 1) : 65536);
}

}  // namespace
}  // namespace fls

using namespace fls;

extern "C" int fls_gpu_selection_indices(const uint64_t* dev_bitmap, uint64_t n_vec, uint16_t* dev_indices,
                                         uint32_t* dev_counts, void* stream) {
    if (n_vec && (!dev_bitmap || !dev_indices)) return fail(FLS_E_ARG, "null argument");
    if (!n_vec) return FLS_OK;
    hipLaunchKernelGGL(k_sel_indices, dim3(waves_grid(n_vec)), dim3(256), 0, (hipStream_t)stream, dev_bitmap, n_vec,
                       dev_indices, dev_counts);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "kernel launch");
}

extern "C" int fls_gpu_selection_and(const uint64_t* dev_a, const uint64_t* dev_b, uint64_t n_vec, uint64_t* dev_out,
                                     uint32_t* dev_counts, void* stream) {
    if (n_vec && (!dev_a || !dev_b || !dev_out)) return fail(FLS_E_ARG, "null argument");
    if (!n_vec) return FLS_OK;
    hipLaunchKernelGGL(k_sel_and, dim3(waves_grid(n_vec)), dim3(256), 0, (hipStream_t)stream, dev_a, dev_b, dev_out,
                       n_vec, dev_counts);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "kernel launch");
}
