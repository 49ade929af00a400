// ref_shim.cpp -- extern "C" entry points over the reference's own compiled code (test infrastructure only).
//
// Linked by oracle/Makefile (target `ref`) against the reference's generated scalar primitives
// (primitives/fls_generated/fallback/scalar_aav_uf1/*_src.cpp) and its vendored FSST
// (src/cor/prm/fsst/libfsst.cpp, fsst_avx512.cpp), compiled from where they lie in the reference tree.
// The result, oracle/_ref/libflsref.so, is used only by tests/ (to pin the oracle and to write golden
// fixtures) and by bench.py's cpu_baseline leg.
#include "fls/cor/prm/fsst/fsst.h"
#include "fls_gen/ffor/ffor.hpp"
#include "fls_gen/pack/pack.hpp"
#include "fls_gen/rsum/rsum.hpp"
#include "fls_gen/transpose/transpose.hpp"
#include "fls_gen/unffor/unffor.hpp"
#include "fls_gen/unpack/unpack.hpp"
#include "fls_gen/unrsum/unrsum.hpp"
#include "fls_gen/untranspose/untranspose.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {
template <typename T>
struct W;
template <>
struct W<uint8_t> {
    static constexpr unsigned bits = 8;
};
template <>
struct W<uint16_t> {
    static constexpr unsigned bits = 16;
};
template <>
struct W<uint32_t> {
    static constexpr unsigned bits = 32;
};
template <>
struct W<uint64_t> {
    static constexpr unsigned bits = 64;
};

template <typename F>
void by_width(unsigned w, F&& f) {
    switch (w) {
    case 8: f(uint8_t {}); break;
    case 16: f(uint16_t {}); break;
    case 32: f(uint32_t {}); break;
    default: f(uint64_t {}); break;
    }
}
} // namespace

extern "C" {

void ref_ffor(unsigned w, const void* in, void* out, unsigned bw, const void* base) {
    by_width(w, [&](auto t) {
        using T = decltype(t);
        generated::ffor::fallback::scalar::ffor(static_cast<const T*>(in), static_cast<T*>(out), (uint8_t)bw,
                                                static_cast<const T*>(base));
    });
}
void ref_unffor(unsigned w, const void* in, void* out, unsigned bw, const void* base) {
    by_width(w, [&](auto t) {
        using T = decltype(t);
        generated::unffor::fallback::scalar::unffor(static_cast<const T*>(in), static_cast<T*>(out), (uint8_t)bw,
                                                    static_cast<const T*>(base));
    });
}
void ref_pack(unsigned w, const void* in, void* out, unsigned bw) {
    by_width(w, [&](auto t) {
        using T = decltype(t);
        generated::pack::fallback::scalar::pack(static_cast<const T*>(in), static_cast<T*>(out), (uint8_t)bw);
    });
}
void ref_unpack(unsigned w, const void* in, void* out, unsigned bw) {
    by_width(w, [&](auto t) {
        using T = decltype(t);
        generated::unpack::fallback::scalar::unpack(static_cast<const T*>(in), static_cast<T*>(out), (uint8_t)bw);
    });
}
void ref_rsum(unsigned w, const void* in, void* out, const void* base) {
    by_width(w, [&](auto t) {
        using T = decltype(t);
        generated::rsum::fallback::scalar::rsum(static_cast<const T*>(in), static_cast<T*>(out),
                                                static_cast<const T*>(base));
    });
}
void ref_unrsum(unsigned w, const void* in, void* out) {
    by_width(w, [&](auto t) {
        using T = decltype(t);
        generated::unrsum::fallback::scalar::unrsum(static_cast<const T*>(in), static_cast<T*>(out));
    });
}
void ref_transpose_i(unsigned w, const void* in, void* out) {
    by_width(w, [&](auto t) {
        using T = decltype(t);
        generated::transpose::fallback::scalar::transpose_i(static_cast<const T*>(in), static_cast<T*>(out));
    });
}
void ref_untranspose_i(unsigned w, const void* in, void* out) {
    by_width(w, [&](auto t) {
        using T = decltype(t);
        generated::untranspose::fallback::scalar::untranspose_i(static_cast<const T*>(in), static_cast<T*>(out));
    });
}
void ref_untranspose_o(unsigned w, const void* in, void* out) {
    by_width(w, [&](auto t) {
        using T = decltype(t);
        generated::untranspose::fallback::scalar::untranspose_o(static_cast<const T*>(in), static_cast<T*>(out));
    });
}

// FSST: build a table over n strings (Arrow offsets), export it, and compress every string with it.
// Returns the exported table length; writes compressed bytes and n+1 compressed offsets.
unsigned ref_fsst_encode(const uint32_t* offsets, const uint8_t* bytes, uint64_t n, uint8_t* table_out,
                         uint8_t* comp_out, uint64_t comp_cap, uint32_t* comp_offsets) {
    std::vector<uint32_t>       len(n);
    std::vector<unsigned char*> ptr(n);
    for (uint64_t i = 0; i < n; ++i) {
        len[i] = offsets[i + 1] - offsets[i];
        ptr[i] = const_cast<unsigned char*>(bytes + offsets[i]);
    }
    fsst_encoder_t* enc = fsst_create(n, len.data(), ptr.data(), 0);
    unsigned        tl  = fsst_export(enc, table_out);
    std::vector<uint32_t>       olen(n);
    std::vector<unsigned char*> optr(n);
    size_t got = fsst_compress(enc, n, len.data(), ptr.data(), comp_cap, comp_out, olen.data(), optr.data());
    comp_offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) comp_offsets[i + 1] = i < got ? comp_offsets[i] + olen[i] : comp_offsets[i];
    fsst_destroy(enc);
    return got == n ? tl : 0;
}
// Decode with the reference's inline fsst_decompress.
uint64_t ref_fsst_decode(const uint8_t* table, const uint8_t* comp, const uint32_t* comp_offsets, uint64_t n,
                         uint8_t* out, uint64_t out_cap, uint32_t* out_offsets) {
    fsst_decoder_t dec;
    if (!fsst_import(&dec, table)) return 0;
    uint64_t o     = 0;
    out_offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) {
        o += fsst_decompress(&dec, comp_offsets[i + 1] - comp_offsets[i], const_cast<uint8_t*>(comp + comp_offsets[i]),
                             out_cap - o, out + o);
        out_offsets[i + 1] = (uint32_t)o;
    }
    return o;
}

// CPU baseline: the reference's per-vector decode of an FFOR stream (unffor + Materializer untranspose_i,
// src/encoder/exp_col_decoder.cpp:13-26, src/encoder/materializer.cpp:13-21) or of the raw primitive
// (untranspose = 0), over n_vec vectors on n_threads threads. The threads are started, and each decodes its slice
// once, before the timed passes; a pass releases every thread through a generation counter and ends when the last
// one has finished its slice. Returns seconds per pass (median of n_iter).
double ref_time_unffor(unsigned w, const void* in, void* out, uint64_t n_vec, unsigned bw, const void* bases,
                       int untranspose, int n_iter, int n_threads) {
    const unsigned    sz = w / 8;
    std::atomic<int>  gen {0}, done {0};
    std::atomic<bool> quit {false};
    auto slice = [&](int k, std::vector<uint8_t>& tmp) {
        const uint64_t v0 = n_vec * k / n_threads, v1 = n_vec * (k + 1) / n_threads;
        for (uint64_t v = v0; v < v1; ++v) {
            const uint8_t* src = static_cast<const uint8_t*>(in) + v * 128 * bw;
            uint8_t*       dst = static_cast<uint8_t*>(out) + v * 1024 * sz;
            const uint8_t* b   = static_cast<const uint8_t*>(bases) + v * sz;
            if (untranspose) {
                ref_unffor(w, src, tmp.data(), bw, b);
                ref_untranspose_i(w, tmp.data(), dst);
            } else {
                ref_unffor(w, src, dst, bw, b);
            }
        }
    };
    std::vector<std::thread> th;
    for (int k = 0; k < n_threads; ++k)
        th.emplace_back([&, k] {
            std::vector<uint8_t> tmp(1024 * 8);
            int                  seen = 0;
            for (;;) {
                gen.wait(seen, std::memory_order_acquire);  // sleeps until the next pass (or the end)
                if (quit.load(std::memory_order_acquire)) return;
                seen = gen.load(std::memory_order_acquire);
                slice(k, tmp);
                if (done.fetch_add(1, std::memory_order_acq_rel) + 1 == n_threads) done.notify_one();
            }
        });
    auto pass = [&] {
        done.store(0, std::memory_order_release);
        gen.fetch_add(1, std::memory_order_acq_rel);
        gen.notify_all();
        for (int d; (d = done.load(std::memory_order_acquire)) < n_threads;) done.wait(d, std::memory_order_acquire);
    };
    pass();  // untimed: every thread has run and touched its slice
    std::vector<double> t(n_iter);
    for (int it = 0; it < n_iter; ++it) {
        const auto t0 = std::chrono::steady_clock::now();
        pass();
        t[it] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    quit.store(true, std::memory_order_release);
    gen.fetch_add(1, std::memory_order_acq_rel);
    gen.notify_all();
    for (auto& x : th) x.join();
    std::sort(t.begin(), t.end());
    return t[n_iter / 2];
}
}
