// k_stream.hip -- raw FFOR / unpack stream kernels, one instantiation per (width, bit width, store mode).
#include <array>

#include "launch.hpp"
#include "page_kernels.hpp"

namespace fls {
namespace {
using StreamFn = void (*)(const uint8_t*, const void*, uint8_t*, uint64_t);

template <int W, bool NT, int... B>
constexpr auto stream_table(std::integer_sequence<int, B...>) {
    return std::array<StreamFn, sizeof...(B)> {&k_stream<W, B, NT>...};
}
template <int W, bool NT>
const auto kTab = stream_table<W, NT>(std::make_integer_sequence<int, W + 1> {});
}  // namespace

hipError_t launch_unffor_stream(unsigned W, unsigned bw, const uint8_t* packed, const void* bases, uint8_t* out,
                                uint64_t n_vec, hipStream_t s) {
    if (bw > W) return hipErrorInvalidValue;
    if (n_vec == 0) return hipSuccess;
    const Tune t = tune_for(FAM_STREAM, W, bw);
    StreamFn   f = nullptr;
    switch (W) {
    case 8: f = (t.nt & 1) ? kTab<8, true>[bw] : kTab<8, false>[bw]; break;
    case 16: f = (t.nt & 1) ? kTab<16, true>[bw] : kTab<16, false>[bw]; break;
    case 32: f = (t.nt & 1) ? kTab<32, true>[bw] : kTab<32, false>[bw]; break;
    case 64: f = (t.nt & 1) ? kTab<64, true>[bw] : kTab<64, false>[bw]; break;
    default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(f, dim3(grid_for(n_vec, t, kTPV)), dim3(t.block), 0, s, packed, bases, out, n_vec);
    return hipGetLastError();
}

}  // namespace fls
