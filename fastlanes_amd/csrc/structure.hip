// structure.hip -- the structural wizard on the GPU: constant, equal and one-to-one mapped columns of a row-group.
//
// Wizard::Spell (src/wizard/wizard.cpp:19-44) plans a row-group before any expression is chosen:
//   constant_check  (:88-93)   a column whose values are all equal (stats is_constant, src/table/attribute.cpp:
//                              101-138: every value equals the previous one) -> CONSTANT;
//   equality_check  (:96-111)  for every pair i < j with j still open: same physical type and equal values
//                              (src/table/rowgroup.cpp:113-156) -> j EQUAL;
//   map_1t1_check   (:125-151) for every open first f with an open second s > f, if f's distinct-value ratio is at
//                              most 0.25 (src/table/rowgroup.cpp:192-217) and each value of f always meets the same
//                              value of s (a hash map from f's values to s's first value, :159-189) -> f DICT +
//                              SUB_EXPRESSION, s FIXME. f stays a candidate for later seconds;
//   expression_check (:114-123) every column left -> ZERO_COPY.
// Values compare as the reference's types do: integers bit for bit, doubles with == (so -0.0 equals 0.0 and a NaN
// equals nothing), strings byte for byte.
//
// GPU form: per column one pass for the constant test and an order-dependent content hash (equality candidates;
// equal hashes are confirmed by a full comparison); per candidate first a lock-free hash set of its distinct values
// (each slot holds 32 hash bits and the row of the value's first insertion, so no false merges: equal hash bits are
// confirmed against that row), which gives both the distinct count and a code per row; then per (first, second)
// pair one pass checking that every row agrees with its code's representative row. The host replays the
// reference's decision order over these results.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../include/fls_gpu.h"
#include "../../include/fls_wizard.h"
#include "gpu_internal.hpp"

namespace fls {
namespace {

enum KeyT { K8, K16, K32, K64, KF64, KSTR };

struct SCol {
    const uint8_t*  data;
    const uint64_t* off;  // strings: n + 1 byte offsets into data
};

// splitmix64 finalizer (Steele, Lea, Flood: "Fast splittable pseudorandom number generators", 2014)
__device__ __forceinline__ uint64_t smix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

template <int KT>
__device__ __forceinline__ uint64_t ival(const SCol& c, uint64_t i) {
    if constexpr (KT == K8) return c.data[i];
    else if constexpr (KT == K16) return reinterpret_cast<const uint16_t*>(c.data)[i];
    else if constexpr (KT == K32) return reinterpret_cast<const uint32_t*>(c.data)[i];
    else return reinterpret_cast<const uint64_t*>(c.data)[i];
}

// Bytes [p, p + n) as one little-endian word (n <= 8), read byte by byte: never past the string.
__device__ __forceinline__ uint64_t tail_word(const uint8_t* p, uint64_t n) {
    uint64_t w = 0;
    for (uint64_t k = 0; k < n; ++k) w |= (uint64_t)p[k] << (8 * k);
    return w;
}

template <int KT>
__device__ __forceinline__ uint64_t key_hash(const SCol& c, uint64_t i) {
    if constexpr (KT == KF64) {
        const double x = reinterpret_cast<const double*>(c.data)[i];
        return smix(x == 0.0 ? 0ull : (uint64_t)__double_as_longlong(x));  // -0.0 and 0.0 hash alike
    } else if constexpr (KT == KSTR) {
        const uint64_t b = c.off[i], e = c.off[i + 1];
        uint64_t       h = smix(e - b);
        uint64_t       p = b;
        for (; p + 8 <= e; p += 8) h = smix(h ^ *reinterpret_cast<const uint64_t __attribute__((aligned(1)))*>(c.data + p));
        if (p < e) h = smix(h ^ tail_word(c.data + p, e - p));
        return h;
    } else {
        return smix(ival<KT>(c, i));
    }
}

template <int KT>
__device__ __forceinline__ bool key_eq(const SCol& a, uint64_t i, const SCol& b, uint64_t j) {
    if constexpr (KT == KF64) {
        return reinterpret_cast<const double*>(a.data)[i] == reinterpret_cast<const double*>(b.data)[j];
    } else if constexpr (KT == KSTR) {
        const uint64_t ab = a.off[i], ae = a.off[i + 1], bb = b.off[j], be = b.off[j + 1];
        if (ae - ab != be - bb) return false;
        for (uint64_t k = 0; k < ae - ab; ++k)
            if (a.data[ab + k] != b.data[bb + k]) return false;
        return true;
    } else {
        return ival<KT>(a, i) == ival<KT>(b, j);
    }
}

#define GRID_STRIDE(i, n) for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

// flag[0] = 1 when some row differs from row 0 (then the column is not constant).
template <int KT>
__global__ void k_not_constant(SCol c, uint64_t n, uint32_t* flag) {
    GRID_STRIDE(i, n) {
        if (i && !key_eq<KT>(c, i, c, 0)) flag[0] = 1u;
    }
}

// Order-dependent content hash (wrapping sum over rows of a mix of the value hash and the row) and, for doubles,
// whether a NaN occurs (a column with a NaN equals no column).
template <int KT>
__global__ void k_content_hash(SCol c, uint64_t n, unsigned long long* hash, uint32_t* nan) {
    uint64_t h = 0;
    bool     has_nan = false;
    GRID_STRIDE(i, n) {
        h += smix(key_hash<KT>(c, i) ^ (i * 0x9e3779b97f4a7c15ull));
        if constexpr (KT == KF64) {
            const double x = reinterpret_cast<const double*>(c.data)[i];
            has_nan |= x != x;
        }
    }
    // one atomic per block: wave sums, then the block's waves through LDS
    __shared__ uint64_t part[16];
    for (int d = 32; d >= 1; d >>= 1) h += __shfl_xor(h, d, 64);
    if (threadIdx.x % 64 == 0) part[threadIdx.x / 64] = h;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (unsigned w = 0; w < blockDim.x / 64; ++w) t += part[w];
        atomicAdd(hash, (unsigned long long)t);
    }
    if (has_nan) nan[0] = 1u;
}

template <int KT>
__global__ void k_columns_differ(SCol a, SCol b, uint64_t n, uint32_t* flag) {
    GRID_STRIDE(i, n) {
        if (!key_eq<KT>(a, i, b, i)) flag[0] = 1u;
    }
}

// Distinct values of a column into an open-addressing set of `cap` (power of two) 64-bit slots: 0 = empty, else
// (32 hash bits << 32) | (row + 1) of the value's first insertion. codes[i] = the slot of row i's value. count =
// slots taken (added once per wave); once it passes `limit` the column is too diverse and `stop` ends the pass.
// cap >= 2 * rows, so the set is never more than half full and every probe sequence ends at an empty or matching
// slot; the cap-step bound is a second guarantee.
template <int KT>
__global__ void k_distinct(SCol c, uint64_t n, unsigned long long* table, uint32_t cap, uint32_t* codes,
                           uint32_t* count, uint32_t limit, uint32_t* stop) {
    GRID_STRIDE(i, n) {
        if (__hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
        const uint64_t           h   = key_hash<KT>(c, i);
        const unsigned long long tag = (h >> 32) << 32;
        uint32_t                 s   = (uint32_t)h & (cap - 1);
        uint32_t                 code = 0xffffffffu;
        for (uint32_t probe = 0; probe < cap; ++probe, s = (s + 1) & (cap - 1)) {
            unsigned long long w = __hip_atomic_load(table + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (w == 0ull) {
                const unsigned long long mine = tag | (i + 1);
                w = atomicCAS(table + s, 0ull, mine);
                if (w == 0ull) {
                    code = s;
                    break;
                }
            }
            if ((w >> 32) == (tag >> 32) && key_eq<KT>(c, i, c, (w & 0xffffffffull) - 1)) {
                code = s;
                break;
            }
        }
        if (code == 0xffffffffu) __hip_atomic_store(stop, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        codes[i] = code;
        // new slots are counted once per wave: the lowest active lane adds the wave's total
        const bool     fresh  = code != 0xffffffffu && (uint32_t)(table[code] & 0xffffffffull) == (uint32_t)(i + 1);
        const uint64_t active = __ballot(1);
        const uint64_t mask   = __ballot(fresh);
        const int      leader = __ffsll((long long)active) - 1;
        if ((int)__lane_id() == leader && mask) {
            const uint32_t add = (uint32_t)__popcll(mask);
            if (atomicAdd(count, add) + add > limit) __hip_atomic_store(stop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// flag[0] = 1 when some row's value of the second column differs from the value at its code's first row.
template <int KT>
__global__ void k_not_mapped(SCol s, uint64_t n, const uint32_t* codes, const unsigned long long* table,
                             uint32_t* flag) {
    GRID_STRIDE(i, n) {
        const uint64_t r = (table[codes[i]] & 0xffffffffull) - 1;
        if (r != i && !key_eq<KT>(s, i, s, r)) flag[0] = 1u;
    }
}

int key_of(uint8_t type) {
    switch (type) {
    case DT_I8: case DT_U8: return K8;
    case DT_I16: case DT_U16: return K16;
    case DT_I32: case DT_U32: return K32;
    case DT_I64: case DT_U64: return K64;
    case DT_DOUBLE: return KF64;
    case DT_STR: return KSTR;
    default: return -1;
    }
}

unsigned grid_for_rows(uint64_t n, unsigned max_blocks = 4096) {
    const uint64_t b = (n + 255) / 256;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, max_blocks));
}

// Runs the statements with KT bound to the key type kt.
#define KEY_SWITCH(kt, ...)                                        \
    switch (kt) {                                                  \
    case K8: { constexpr int KT = K8; __VA_ARGS__; } break;       \
    case K16: { constexpr int KT = K16; __VA_ARGS__; } break;     \
    case K32: { constexpr int KT = K32; __VA_ARGS__; } break;     \
    case K64: { constexpr int KT = K64; __VA_ARGS__; } break;     \
    case KF64: { constexpr int KT = KF64; __VA_ARGS__; } break;   \
    default: { constexpr int KT = KSTR; __VA_ARGS__; } break;     \
    }

struct Flags {  // per column, device
    uint32_t           not_const, nan, differ, pad;
    unsigned long long hash;
};

}  // namespace
}  // namespace fls

using namespace fls;

extern "C" int fls_wizard_spell(uint32_t n_columns, const struct fls_spell_column* columns, uint64_t n_rows,
                                int32_t* out_kind, int32_t* out_ref, void* stream) {
    if (n_columns && (!columns || !out_kind || !out_ref)) return fail(FLS_E_ARG, "null argument");
    if (n_rows > (1ull << 30)) return fail(FLS_E_ARG, "a row-group holds at most 2^30 rows");
    std::vector<int> kt(n_columns);
    for (uint32_t c = 0; c < n_columns; ++c) {
        kt[c] = key_of(columns[c].type);
        if (kt[c] < 0) return fail(FLS_E_UNSUPPORTED, "column " + std::to_string(c) + ": unsupported data type");
        if (n_rows && !columns[c].values && !(kt[c] == KSTR && columns[c].offsets))
            return fail(FLS_E_ARG, "column " + std::to_string(c) + ": null values");
        if (kt[c] == KSTR && !columns[c].offsets) return fail(FLS_E_ARG, "string column without offsets");
        out_kind[c] = FLS_SPELL_EXPRESSION, out_ref[c] = -1;
    }
    if (!n_columns) return FLS_OK;
    hipStream_t s = (hipStream_t)stream;
    hipError_t  e;
    auto        col = [&](uint32_t c) { return SCol {(const uint8_t*)columns[c].values, columns[c].offsets}; };
    const uint64_t n = n_rows;

    // 1) constant test and content hash of every column
    Flags* d_flags = nullptr;
    if ((e = hipMallocAsync((void**)&d_flags, sizeof(Flags) * n_columns, s)) != hipSuccess) return hip_fail(e, "hipMallocAsync");
    std::vector<Flags> flags(n_columns);
    auto finish = [&](int rc) {
        (void)hipFreeAsync(d_flags, s);
        return rc;
    };
    if ((e = hipMemsetAsync(d_flags, 0, sizeof(Flags) * n_columns, s)) != hipSuccess) return finish(hip_fail(e, "hipMemsetAsync"));
    for (uint32_t c = 0; c < n_columns && n; ++c) {
        const SCol sc = col(c);
        const dim3 g(grid_for_rows(n)), b(256);
        KEY_SWITCH(kt[c], {
            hipLaunchKernelGGL((k_not_constant<KT>), g, b, 0, s, sc, n, &d_flags[c].not_const);
            hipLaunchKernelGGL((k_content_hash<KT>), dim3(grid_for_rows(n, 1024)), b, 0, s, sc, n, &d_flags[c].hash,
                               &d_flags[c].nan);
        });
    }
    if ((e = hipGetLastError()) != hipSuccess) return finish(hip_fail(e, "kernel launch"));
    if ((e = hipMemcpyAsync(flags.data(), d_flags, sizeof(Flags) * n_columns, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return finish(hip_fail(e, "constant test"));
    std::vector<bool> det(n_columns, false);
    for (uint32_t c = 0; c < n_columns; ++c)
        if (!flags[c].not_const) out_kind[c] = FLS_SPELL_CONSTANT, det[c] = true;

    // 2) equality: j takes the first i < j of its type with equal values (a constant i has no open equal j)
    for (uint32_t j = 0; j < n_columns; ++j) {
        if (det[j] || flags[j].nan) continue;
        for (uint32_t i = 0; i < j; ++i) {
            if (columns[i].type != columns[j].type || flags[i].nan || flags[i].hash != flags[j].hash || !flags[i].not_const)
                continue;
            uint32_t differ = 0;
            if ((e = hipMemsetAsync(&d_flags[j].differ, 0, 4, s)) != hipSuccess) return finish(hip_fail(e, "hipMemsetAsync"));
            const SCol a = col(i), b2 = col(j);
            KEY_SWITCH(kt[j], hipLaunchKernelGGL((k_columns_differ<KT>), dim3(grid_for_rows(n)), dim3(256), 0, s, a, b2, n,
                                                 &d_flags[j].differ));
            if ((e = hipMemcpyAsync(&differ, &d_flags[j].differ, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
                (e = hipStreamSynchronize(s)) != hipSuccess)
                return finish(hip_fail(e, "equality test"));
            if (!differ) {
                out_kind[j] = FLS_SPELL_EQUAL, out_ref[j] = (int32_t)i, det[j] = true;
                break;
            }
        }
    }

    // 3) one-to-one maps, in the reference's order
    const uint32_t limit = (uint32_t)(n / 4);  // distinct / rows <= 0.25
    uint32_t       cap   = 64;
    while (cap < 2ull * n) cap <<= 1;
    unsigned long long* d_table = nullptr;
    uint32_t*           d_codes = nullptr;
    uint32_t*           d_cnt   = nullptr;  // [0] count, [1] stop, [2..] per-second flags
    auto                finish2 = [&](int rc) {
        if (d_table) (void)hipFreeAsync(d_table, s);
        if (d_codes) (void)hipFreeAsync(d_codes, s);
        if (d_cnt) (void)hipFreeAsync(d_cnt, s);
        return finish(rc);
    };
    for (uint32_t f = 0; f < n_columns; ++f) {
        if (det[f] || !n) continue;
        std::vector<uint32_t> seconds;
        for (uint32_t t = f + 1; t < n_columns; ++t)
            if (!det[t]) seconds.push_back(t);
        if (seconds.empty()) continue;
        if (!d_table) {
            if ((e = hipMallocAsync((void**)&d_table, 8ull * cap, s)) != hipSuccess ||
                (e = hipMallocAsync((void**)&d_codes, 4ull * n, s)) != hipSuccess ||
                (e = hipMallocAsync((void**)&d_cnt, 4ull * (2 + n_columns), s)) != hipSuccess)
                return finish2(hip_fail(e, "hipMallocAsync"));
        }
        if ((e = hipMemsetAsync(d_table, 0, 8ull * cap, s)) != hipSuccess ||
            (e = hipMemsetAsync(d_cnt, 0, 4ull * (2 + n_columns), s)) != hipSuccess)
            return finish2(hip_fail(e, "hipMemsetAsync"));
        const SCol fc = col(f);
        // 2^18 threads: a column with more than n / 4 distinct values stops after about one row per thread
        KEY_SWITCH(kt[f], hipLaunchKernelGGL((k_distinct<KT>), dim3(grid_for_rows(n, 1024)), dim3(256), 0, s, fc, n,
                                             d_table, cap, d_codes, d_cnt, limit, d_cnt + 1));
        uint32_t hc[2];
        if ((e = hipGetLastError()) != hipSuccess ||
            (e = hipMemcpyAsync(hc, d_cnt, 8, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return finish2(hip_fail(e, "distinct values"));
        if (hc[1] || hc[0] > limit) continue;  // too many distinct values for a dictionary
        for (uint32_t t : seconds) {
            const SCol sc = col(t);
            KEY_SWITCH(kt[t], hipLaunchKernelGGL((k_not_mapped<KT>), dim3(grid_for_rows(n)), dim3(256), 0, s, sc, n, d_codes,
                                                 d_table, d_cnt + 2 + t));
        }
        std::vector<uint32_t> nm(n_columns, 1);
        if ((e = hipGetLastError()) != hipSuccess ||
            (e = hipMemcpyAsync(nm.data(), d_cnt + 2, 4ull * n_columns, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return finish2(hip_fail(e, "map test"));
        for (uint32_t t : seconds) {
            if (nm[t]) continue;
            if (out_kind[f] != FLS_SPELL_MAP_SOURCE) out_kind[f] = FLS_SPELL_MAP_SOURCE, out_ref[f] = (int32_t)t;
            det[f]      = true;
            out_kind[t] = FLS_SPELL_MAP_TARGET, out_ref[t] = (int32_t)f, det[t] = true;
        }
    }
    int rc = finish2(FLS_OK);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "release");
    return rc;
}
