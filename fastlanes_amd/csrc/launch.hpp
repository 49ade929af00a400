// launch.hpp -- host-side launchers for the decode kernels (one translation unit per kernel family).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fls {

struct HzlChunk;

hipError_t launch_unffor_stream(unsigned W, unsigned bw, const uint8_t* packed, const void* bases, uint8_t* out,
                                uint64_t n_vec, hipStream_t s);
hipError_t launch_page_int(unsigned W, int mode, const uint8_t* page, uint32_t arr_c, uint32_t vec_c, uint8_t* out,
                           hipStream_t s);
hipError_t launch_page_rle(unsigned W, int var, const uint8_t* page, uint32_t arr_c, uint32_t vec_c, uint8_t* out,
                           uint32_t* err, hipStream_t s);
hipError_t launch_page_dict(unsigned W, int var, const uint8_t* page, uint32_t arr_c, uint32_t vec_c,
                            const void* dict, uint32_t dict_n, uint8_t* out, uint32_t* err, hipStream_t s);
hipError_t launch_hzl(unsigned DW, const uint8_t* enc, const HzlChunk* chunks, uint32_t n_chunks, uint8_t* dict,
                      hipStream_t s);

}  // namespace fls
