// k_stream.hip -- raw FFOR / unpack stream kernels, one instantiation per (width, bit width).
#include "launch.hpp"
#include "page_kernels.hpp"

namespace fls {
namespace {
using StreamFn = void (*)(const uint8_t*, const void*, uint8_t*, uint64_t);

template <int W, int... B>
constexpr auto stream_table(std::integer_sequence<int, B...>) {
    return std::array<StreamFn, sizeof...(B)> {&k_unffor_stream<W, B>...};
}
const auto kS8  = stream_table<8>(std::make_integer_sequence<int, 9> {});
const auto kS16 = stream_table<16>(std::make_integer_sequence<int, 17> {});
const auto kS32 = stream_table<32>(std::make_integer_sequence<int, 33> {});
const auto kS64 = stream_table<64>(std::make_integer_sequence<int, 65> {});
}  // namespace

hipError_t launch_unffor_stream(unsigned W, unsigned bw, const uint8_t* packed, const void* bases, uint8_t* out,
                                uint64_t n_vec, hipStream_t s) {
    if (bw > W) return hipErrorInvalidValue;
    StreamFn f = W == 8 ? kS8[bw] : W == 16 ? kS16[bw] : W == 32 ? kS32[bw] : kS64[bw];
    const unsigned vpb = W == 64 ? Geo<64>::VPB : Geo<32>::VPB;
    if (n_vec == 0) return hipSuccess;
    const uint64_t blocks = (n_vec + vpb - 1) / vpb;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(f, dim3((unsigned)blocks), dim3(kBlock), 0, s, packed, bases, out, n_vec);
    return hipGetLastError();
}

}  // namespace fls
