// k_page.hip -- page decode kernels (UNCOMPRESSED / FFOR / DELTA / RLE / DICT) and dictionary payload kernels.
#include "launch.hpp"
#include "page_kernels.hpp"

namespace fls {

template <int W>
static hipError_t page_int_w(int mode, const uint8_t* page, uint32_t arr_c, uint32_t vec_c, uint8_t* out,
                             hipStream_t s) {
    const unsigned blocks = (vec_c + Geo<W>::VPB - 1) / Geo<W>::VPB;
    if (mode == 0) hipLaunchKernelGGL((k_page_int<W, 0>), dim3(blocks), dim3(kBlock), 0, s, page, arr_c, vec_c, out);
    else if (mode == 1) hipLaunchKernelGGL((k_page_int<W, 1>), dim3(blocks), dim3(kBlock), 0, s, page, arr_c, vec_c, out);
    else hipLaunchKernelGGL((k_page_int<W, 2>), dim3(blocks), dim3(kBlock), 0, s, page, arr_c, vec_c, out);
    return hipGetLastError();
}

hipError_t launch_page_int(unsigned W, int mode, const uint8_t* page, uint32_t arr_c, uint32_t vec_c, uint8_t* out,
                           hipStream_t s) {
    if (!vec_c) return hipSuccess;
    switch (W) {
    case 8: return page_int_w<8>(mode, page, arr_c, vec_c, out, s);
    case 16: return page_int_w<16>(mode, page, arr_c, vec_c, out, s);
    case 32: return page_int_w<32>(mode, page, arr_c, vec_c, out, s);
    case 64: return page_int_w<64>(mode, page, arr_c, vec_c, out, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
static hipError_t page_rle_t(int var, const uint8_t* page, uint32_t arr_c, uint32_t vec_c, uint8_t* out,
                             uint32_t* err, hipStream_t s) {
    const unsigned blocks = (vec_c + Geo<16>::VPB - 1) / Geo<16>::VPB;
    T*             o      = reinterpret_cast<T*>(out);
    if (var == 1) hipLaunchKernelGGL((k_page_rle<T, 1>), dim3(blocks), dim3(kBlock), 0, s, page, arr_c, vec_c, o, err);
    else hipLaunchKernelGGL((k_page_rle<T, 2>), dim3(blocks), dim3(kBlock), 0, s, page, arr_c, vec_c, o, err);
    return hipGetLastError();
}

hipError_t launch_page_rle(unsigned W, int var, const uint8_t* page, uint32_t arr_c, uint32_t vec_c, uint8_t* out,
                           uint32_t* err, hipStream_t s) {
    if (!vec_c) return hipSuccess;
    switch (W) {
    case 8: return page_rle_t<uint8_t>(var, page, arr_c, vec_c, out, err, s);
    case 16: return page_rle_t<uint16_t>(var, page, arr_c, vec_c, out, err, s);
    case 32: return page_rle_t<uint32_t>(var, page, arr_c, vec_c, out, err, s);
    case 64: return page_rle_t<uint64_t>(var, page, arr_c, vec_c, out, err, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename T, typename D>
static hipError_t page_dict_t(int var, const uint8_t* page, uint32_t arr_c, uint32_t vec_c, const void* dict,
                              uint32_t dict_n, uint8_t* out, uint32_t* err, hipStream_t s) {
    const unsigned blocks = (vec_c + Geo<32>::VPB - 1) / Geo<32>::VPB;
    T*             o      = reinterpret_cast<T*>(out);
    const D*       d      = reinterpret_cast<const D*>(dict);
    if (var == 1)
        hipLaunchKernelGGL((k_page_dict<T, D, 1>), dim3(blocks), dim3(kBlock), 0, s, page, arr_c, vec_c, d, dict_n, o, err);
    else if (var == 2)
        hipLaunchKernelGGL((k_page_dict<T, D, 2>), dim3(blocks), dim3(kBlock), 0, s, page, arr_c, vec_c, d, dict_n, o, err);
    else
        hipLaunchKernelGGL((k_page_dict<T, D, 7>), dim3(blocks), dim3(kBlock), 0, s, page, arr_c, vec_c, d, dict_n, o, err);
    return hipGetLastError();
}

hipError_t launch_page_dict(unsigned W, int var, const uint8_t* page, uint32_t arr_c, uint32_t vec_c,
                            const void* dict, uint32_t dict_n, uint8_t* out, uint32_t* err, hipStream_t s) {
    if (!vec_c) return hipSuccess;
    switch (W) {
    case 8: return page_dict_t<uint8_t, uint32_t>(var, page, arr_c, vec_c, dict, dict_n, out, err, s);
    case 16: return page_dict_t<uint16_t, uint32_t>(var, page, arr_c, vec_c, dict, dict_n, out, err, s);
    case 32: return page_dict_t<uint32_t, uint32_t>(var, page, arr_c, vec_c, dict, dict_n, out, err, s);
    case 64: return page_dict_t<uint64_t, uint64_t>(var, page, arr_c, vec_c, dict, dict_n, out, err, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_hzl(unsigned DW, const uint8_t* enc, const HzlChunk* chunks, uint32_t n_chunks, uint8_t* dict,
                      hipStream_t s) {
    if (!n_chunks) return hipSuccess;
    const unsigned vpb    = DW == 64 ? Geo<64>::VPB : Geo<32>::VPB;
    const unsigned blocks = (n_chunks + vpb - 1) / vpb;
    if (DW == 32) {
        hipLaunchKernelGGL((k_hzl_chunks<32>), dim3(blocks), dim3(kBlock), 0, s, enc, chunks, n_chunks, dict);
        hipLaunchKernelGGL((k_hzl_exceptions<32>), dim3(n_chunks < 1024 ? n_chunks : 1024), dim3(kBlock), 0, s, enc,
                           chunks, n_chunks, dict);
    } else {
        hipLaunchKernelGGL((k_hzl_chunks<64>), dim3(blocks), dim3(kBlock), 0, s, enc, chunks, n_chunks, dict);
        hipLaunchKernelGGL((k_hzl_exceptions<64>), dim3(n_chunks < 1024 ? n_chunks : 1024), dim3(kBlock), 0, s, enc,
                           chunks, n_chunks, dict);
    }
    return hipGetLastError();
}

}  // namespace fls
