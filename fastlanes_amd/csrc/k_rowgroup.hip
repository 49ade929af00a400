// k_rowgroup.hip -- one persistent launch that decodes a whole row-group from a device-side work list.
//
// The host cuts every page of every column into tiles of consecutive vectors (RgTile). Single-wave blocks take
// tiles from an atomic counter until the list is exhausted, decode each tile with the same per-vector bodies the
// per-page kernels use, stage the wave's vectors in LDS and write whole lines to HBM. Tiles of one kind are
// contiguous in the list so consecutive tiles on a compute unit run the same code path. Kinds outside the fused
// set are decoded by their per-page kernels (rowgroup.hip).
#include "../../include/fls_gpu.h"
#include "gpu_internal.hpp"
#include "page_kernels.hpp"
#include "rowgroup.hpp"

namespace fls {

namespace {

template <unsigned TPV, unsigned VB, typename F>
__device__ __forceinline__ void tile_loop(const RgTile& tl, uint8_t* out, F&& body) {
    // the kernel's dynamic LDS, declared here so that accesses keep the LDS address space inside each function
    extern __shared__ __attribute__((aligned(16))) uint8_t rg_lds[];
    using WS           = WaveStage<TPV, VB>;
    uint8_t* const wl   = rg_lds;
    const unsigned lane = threadIdx.x % 64, t = lane % TPV, slot = lane / TPV;
    const LdsSink  sink {wl + slot * WS::PVB};
    const uint64_t end = (uint64_t)tl.v0 + tl.nv;
    for (uint64_t v0 = tl.v0; v0 < end; v0 += WS::VPW) {
        const uint64_t v = v0 + slot;
        if (v < end) body(v, t, sink);
        wave_sync();
        const uint64_t left = end - v0;
        wave_copy_out<TPV, VB, true>(wl, out + v0 * VB, left < WS::VPW ? (uint32_t)left : WS::VPW, lane);
        wave_sync();
    }
}

// Each kind's tile loop is a separate (not inlined) function: the kernel itself is a small dispatch loop, so
// the compiler schedules and allocates registers per kind instead of over one function holding every body.
template <int W, int MODE>
__device__ __noinline__ void run_int(const RgTile* tp, const uint8_t* enc, uint8_t* out) {
    const RgTile   tl   = *tp;
    const uint8_t* page = enc + tl.page;
    constexpr unsigned TPV = lane_tpv<W>(), VB = 128 * W;
    tile_loop<TPV, VB>(tl, out + tl.out,
                       [&](uint64_t v, unsigned t, const LdsSink& s) { int_page_vector<W, MODE>(page, tl.arr_c, v, t, s); });
}

template <typename T>
__device__ __noinline__ void run_rle(const RgTile* tp, const uint8_t* enc, uint8_t* out, uint32_t* err) {
    const RgTile   tl   = *tp;
    const uint8_t* page = enc + tl.page;
    tile_loop<32, 1024 * sizeof(T)>(tl, out + tl.out, [&](uint64_t v, unsigned t, const LdsSink& s) {
        rle_page_vector<T, 2>(page, tl.arr_c, tl.vec_c, v, t, s, err);
    });
}

template <typename T>
__device__ __noinline__ void run_dict(const RgTile* tp, const uint8_t* enc, uint8_t* out, uint32_t* err) {
    using D             = typename uint_of<(sizeof(T) < 4 ? 32 : 8 * sizeof(T))>::type;
    const RgTile   tl   = *tp;
    const uint8_t* page = enc + tl.page;
    const D*       dict = reinterpret_cast<const D*>(tl.dict);
    if (tl.dict_n == 0) return;
    tile_loop<32, 1024 * sizeof(T)>(tl, out + tl.out, [&](uint64_t v, unsigned t, const LdsSink& s) {
        dict_page_vector<T, D, 2>(page, tl.arr_c, v, t, dict, tl.dict_n, s, err);
    });
}

}  // namespace

__global__ __launch_bounds__(64) void k_rowgroup(const uint8_t* __restrict__ enc, uint8_t* __restrict__ out,
                                                 const RgTile* __restrict__ tiles, uint32_t n_tiles,
                                                 uint32_t* __restrict__ counter, uint32_t* err) {
    for (;;) {
        uint32_t id = 0;
        if (threadIdx.x == 0) id = atomicAdd(counter, 1u);
        id = __shfl(id, 0, 64);
        if (id >= n_tiles) break;
        const RgTile* tp = tiles + id;
        switch (tp->kind) {
        case RG_FFOR8: run_int<8, 1>(tp, enc, out); break;
        case RG_FFOR16: run_int<16, 1>(tp, enc, out); break;
        case RG_FFOR32: run_int<32, 1>(tp, enc, out); break;
        case RG_FFOR64: run_int<64, 1>(tp, enc, out); break;
        case RG_DELTA8: run_int<8, 2>(tp, enc, out); break;
        case RG_DELTA16: run_int<16, 2>(tp, enc, out); break;
        case RG_DELTA32: run_int<32, 2>(tp, enc, out); break;
        case RG_DELTA64: run_int<64, 2>(tp, enc, out); break;
        case RG_UNC8: run_int<8, 0>(tp, enc, out); break;
        case RG_UNC16: run_int<16, 0>(tp, enc, out); break;
        case RG_UNC32: run_int<32, 0>(tp, enc, out); break;
        case RG_UNC64: run_int<64, 0>(tp, enc, out); break;
        case RG_RLE8: run_rle<uint8_t>(tp, enc, out, err); break;
        case RG_RLE16: run_rle<uint16_t>(tp, enc, out, err); break;
        case RG_RLE32: run_rle<uint32_t>(tp, enc, out, err); break;
        case RG_RLE64: run_rle<uint64_t>(tp, enc, out, err); break;
        case RG_DICT8: run_dict<uint8_t>(tp, enc, out, err); break;
        case RG_DICT16: run_dict<uint16_t>(tp, enc, out, err); break;
        case RG_DICT32: run_dict<uint32_t>(tp, enc, out, err); break;
        case RG_DICT64: run_dict<uint64_t>(tp, enc, out, err); break;
        default: atomicOr(err, 8u); break;
        }
    }
}

__global__ void k_rowgroup_reset(uint32_t* counter) { *counter = 0u; }

unsigned rg_stage_bytes(int kind) {
    switch (kind) {
    case RG_FFOR8: case RG_DELTA8: case RG_UNC8: return WaveStage<32, 1024>::BYTES;
    case RG_FFOR16: case RG_DELTA16: case RG_UNC16: return WaveStage<32, 2048>::BYTES;
    case RG_FFOR32: case RG_DELTA32: case RG_UNC32: return WaveStage<32, 4096>::BYTES;
    case RG_FFOR64: case RG_DELTA64: case RG_UNC64: return WaveStage<16, 8192>::BYTES;
    case RG_RLE8: case RG_DICT8: return WaveStage<32, 1024>::BYTES;
    case RG_RLE16: case RG_DICT16: return WaveStage<32, 2048>::BYTES;
    case RG_RLE32: case RG_DICT32: return WaveStage<32, 4096>::BYTES;
    case RG_RLE64: case RG_DICT64: return WaveStage<32, 8192>::BYTES;
    default: return 0;
    }
}

hipError_t launch_rowgroup(const uint8_t* enc, uint8_t* out, const RgTile* tiles, uint32_t n_tiles, uint32_t* counter,
                           uint32_t* err, unsigned lds, unsigned blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_rowgroup_reset, dim3(1), dim3(1), 0, s, counter);
    hipLaunchKernelGGL(k_rowgroup, dim3(blocks), dim3(64), lds, s, enc, out, tiles, n_tiles, counter, err);
    return hipGetLastError();
}

}  // namespace fls
