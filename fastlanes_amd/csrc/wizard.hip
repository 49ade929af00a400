// wizard.hip -- encoding selection kernels: exhaustive trial sizing and the learned selector's sample statistics.
//
// The reference chooses an expression by encoding a column with every expression of its pool and keeping the
// smallest result (ExpColEncoder<PT>::encode_all / AllExpEncodedCol::GetBest, src/encoder/exp_col_encoder.cpp:19-28,
// 117-129). Here the unit is a 65536-value int64 row-group (64 vectors = one page per candidate), and every
// candidate's exact encoded size is computed in closed form from per-vector statistics -- the page layout of
// src/cor/lyt/page/page.cpp:120-130, bit widths of ffor_analyze (src/cor/prm/ffor/ffor_analyze.cpp:78-102, exact bit
// length), the RLE split of src/cor/lyt/vec.cpp:50-71, insertion-order dictionaries (src/cor/lyt/dic/dic_fac.cpp:
// 259-330) and the HZL dictionary payload (src/cor/eng/hzl_cmp.cpp:43-176) -- without producing encoded bytes.
//
// Delta rule used throughout: the transposed lane order of a W-bit type visits aligned W-value segments of the
// original order one after another, so the lane deltas (unrsum) are x[o] - x[o-1] inside each segment and 0 at a
// segment start (64-value segments for int64 values, 32 for u32 dictionary codes, 16 for u16 RLE positions).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/fls_gpu.h"
#include "../../include/fls_wizard.h"
#include "gpu_internal.hpp"

namespace fls {
namespace {

constexpr unsigned kRG = 65536, kVPR = 64, kK = 8;
constexpr uint64_t kEmpty = ~0ull;
constexpr unsigned kHtBits = 17, kHt = 1u << kHtBits;      // exhaustive table: 131072 slots (load <= 0.5)
constexpr unsigned kSampleV = 4, kSample = kSampleV * 1024;  // learned selector's sample
constexpr unsigned kSampleVec[kSampleV] = {0, 21, 42, 63};
constexpr uint8_t kCandExpr[kK] = {0, 1, 11, 6, 7, 16, 17, 22};
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t bitlen64(uint64_t d) { return d ? 64u - (uint32_t)__clzll((long long)d) : 0u; }

// splitmix64 finalizer (Steele, Lea, Flood: "Fast splittable pseudorandom number generators", 2014)
__device__ __forceinline__ uint32_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return (uint32_t)(z ^ (z >> 31));
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
    for (int s = 32; s >= 1; s >>= 1) {
        T o = __shfl_xor(v, s, 64);
        v   = o < v ? o : v;
    }
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
    for (int s = 32; s >= 1; s >>= 1) {
        T o = __shfl_xor(v, s, 64);
        v   = o > v ? o : v;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
    return v;
}

// Statistics of one 1024-value vector, computed by one wave (lane l owns values 16l .. 16l+15).
struct VecStats {
    uint32_t bwf, bwd, runs, bwp, eq, nondec, inc;
    uint64_t umin, umax;
    int64_t  smin, smax;
};

__device__ VecStats vector_stats(const uint64_t* __restrict__ x) {
    const unsigned lane = threadIdx.x % 64;
    uint64_t       v[16];
    const v4u*     p = reinterpret_cast<const v4u*>(x + 16 * lane);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const v4u q  = p[i];
        v[2 * i]     = (uint64_t)q[0] | ((uint64_t)q[1] << 32);
        v[2 * i + 1] = (uint64_t)q[2] | ((uint64_t)q[3] << 32);
    }
    uint64_t prev = __shfl_up(v[15], 1, 64);
    uint64_t mn = kEmpty, mx = 0, dmn = kEmpty, dmx = 0;
    int64_t  smn = INT64_MAX, smx = INT64_MIN;
    uint32_t runs = 0, inner = 0, eq = 0, nondec = 0, inc = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const unsigned o  = 16 * lane + i;
        const uint64_t pv = i ? v[i - 1] : prev;
        mn = v[i] < mn ? v[i] : mn;
        mx = v[i] > mx ? v[i] : mx;
        smn = (int64_t)v[i] < smn ? (int64_t)v[i] : smn;
        smx = (int64_t)v[i] > smx ? (int64_t)v[i] : smx;
        const uint64_t d = (o % 64) ? v[i] - pv : 0ull;
        dmn = d < dmn ? d : dmn;
        dmx = d > dmx ? d : dmx;
        const bool has_prev = o != 0;
        const bool chg      = !has_prev || v[i] != pv;
        runs += chg;
        inner |= (o % 16) != 0 && chg;
        eq += has_prev && v[i] == pv;
        nondec += has_prev && (int64_t)v[i] >= (int64_t)pv;
        inc += has_prev && (int64_t)v[i] > (int64_t)pv;
    }
    VecStats s;
    s.umin = wave_min(mn), s.umax = wave_max(mx);
    s.smin = wave_min(smn), s.smax = wave_max(smx);
    s.bwf  = bitlen64(s.umax - s.umin);
    s.bwd  = bitlen64(wave_max(dmx) - wave_min(dmn));
    s.runs = wave_sum(runs), s.bwp = wave_sum(inner) != 0;
    s.eq = wave_sum(eq), s.nondec = wave_sum(nondec), s.inc = wave_sum(inc);
    return s;
}

// Insert v (first seen at position o) into an open-addressing table whose empty keys are kEmpty; the value kEmpty
// itself lives in the extra slot `mask + 1`. Returns the slot (never fails: callers size tables above their input).
template <typename KeyPtr, typename FirstPtr>
__device__ __forceinline__ uint32_t ht_insert(KeyPtr keys, FirstPtr first, uint32_t mask, uint64_t v, uint32_t o) {
    if (v == kEmpty) {
        atomicMin(&first[mask + 1], o);
        return mask + 1;
    }
    uint32_t h = mix64(v) & mask;
    for (uint32_t p = 0; p <= mask; ++p) {
        const uint64_t old = atomicCAS((unsigned long long*)&keys[h], (unsigned long long)kEmpty, (unsigned long long)v);
        if (old == kEmpty || old == v) {
            atomicMin(&first[h], o);
            return h;
        }
        h = (h + 1) & mask;
    }
    return mask + 1;  // unreachable: the table holds more slots than distinct inputs
}

__device__ __forceinline__ uint32_t rank_of(const uint32_t* bitmap, const uint32_t* base, uint32_t f) {
    return base[f >> 5] + (uint32_t)__popc(bitmap[f >> 5] & ((1u << (f & 31)) - 1u));
}

// exclusive scan of popcounts of `words` bitmap words into base[]; returns the total. All threads of the block.
__device__ uint32_t block_popcount_scan(const uint32_t* bitmap, uint32_t* base, unsigned words, uint32_t* tmp) {
    const unsigned t = threadIdx.x, nt = blockDim.x, per = (words + nt - 1) / nt;
    uint32_t       s = 0;
    for (unsigned i = 0; i < per; ++i) {
        const unsigned w = t * per + i;
        if (w < words) s += (uint32_t)__popc(bitmap[w]);
    }
    // inclusive scan of s over threads (wave shuffles + per-wave totals)
    const unsigned lane = t % 64, wave = t / 64, nw = nt / 64;
    uint32_t       inc  = s;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= (unsigned)d) inc += o;
    }
    if (lane == 63) tmp[wave] = inc;
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (unsigned w = 0; w < nw; ++w) {
            const uint32_t x = tmp[w];
            tmp[w]           = acc;
            acc += x;
        }
        tmp[nw] = acc;
    }
    __syncthreads();
    uint32_t run = tmp[wave] + inc - s;
    for (unsigned i = 0; i < per; ++i) {
        const unsigned w = t * per + i;
        if (w < words) base[w] = run, run += (uint32_t)__popc(bitmap[w]);
    }
    const uint32_t total = tmp[nw];
    __syncthreads();
    return total;
}

// ---------------------------------------------------------------------------------------------------------------
// Exhaustive sizing: one 1024-thread block per row-group (grid-stride); per block a global hash table region.
// ---------------------------------------------------------------------------------------------------------------
struct SizerLds {
    uint32_t bitmap[kRG / 32];
    uint32_t base[kRG / 32];
    uint32_t bwf[kVPR], bwd[kVPR], runs[kVPR], bwp[kVPR];
    uint32_t cmin[kVPR], cmax[kVPR], cdmin[kVPR], cdmax[kVPR];
    unsigned long long dmin[kVPR], dmax[kVPR];  // dictionary chunk min / max (chunks of 1024 codes)
    uint32_t tmp[32];
};

__global__ __launch_bounds__(1024) void k_wizard_sizes(const uint64_t* __restrict__ values, uint64_t n_rg,
                                                       uint64_t* __restrict__ tables, uint32_t* __restrict__ firsts_all,
                                                       uint32_t* __restrict__ slots_all, uint64_t* __restrict__ sizes) {
    __shared__ SizerLds L;
    const unsigned t = threadIdx.x, lane = t % 64, wave = t / 64;
    uint64_t* keys   = tables + (uint64_t)blockIdx.x * (kHt + 64);  // kHt keys + padding
    uint32_t* first  = firsts_all + (uint64_t)blockIdx.x * (2 * kHt + 64);
    uint32_t* fpos   = first + kHt + 32;                               // first occurrence per position (kRG)
    uint32_t* slot   = slots_all + (uint64_t)blockIdx.x * kRG;
    for (uint64_t r = blockIdx.x; r < n_rg; r += gridDim.x) {
        const uint64_t* x = values + r * kRG;
        for (unsigned i = t; i < kRG / 32; i += blockDim.x) L.bitmap[i] = 0;
        if (t < kVPR) {
            L.cmin[t] = ~0u, L.cmax[t] = 0, L.cdmin[t] = ~0u, L.cdmax[t] = 0;
            L.dmin[t] = kEmpty, L.dmax[t] = 0;
        }
        // per-vector statistics
        for (unsigned v = wave; v < kVPR; v += blockDim.x / 64) {
            const VecStats s = vector_stats(x + v * 1024);
            if (lane == 0) L.bwf[v] = s.bwf, L.bwd[v] = s.bwd, L.runs[v] = s.runs, L.bwp[v] = s.bwp;
        }
        // dictionary: insert every value, remember its slot
        for (unsigned o = t; o < kRG; o += blockDim.x) slot[o] = ht_insert(keys, first, kHt - 1, x[o], o);
        __threadfence();
        __syncthreads();
        for (unsigned o = t; o < kRG; o += blockDim.x) {
            const uint32_t f = __hip_atomic_load(&first[slot[o]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            fpos[o]          = f;
            if (f == o) atomicOr(&L.bitmap[o >> 5], 1u << (o & 31));
        }
        __threadfence();
        __syncthreads();
        const uint32_t D = block_popcount_scan(L.bitmap, L.base, kRG / 32, L.tmp);
        // codes: per vector min/max and 32-segment delta min/max; dictionary chunk min/max
        for (unsigned o = t; o < kRG; o += blockDim.x) {
            const uint32_t f  = fpos[o];
            const uint32_t c  = rank_of(L.bitmap, L.base, f);
            const uint32_t cp = (o % 32) ? rank_of(L.bitmap, L.base, fpos[o - 1]) : c;
            const uint32_t d  = c - cp;  // 0 at a segment start
            const uint32_t cmn = wave_min(c), cmx = wave_max(c), dmn = wave_min(d), dmx = wave_max(d);
            if (lane == 0) {
                const unsigned v = o >> 10;  // the whole wave lies in one vector
                atomicMin(&L.cmin[v], cmn), atomicMax(&L.cmax[v], cmx);
                atomicMin(&L.cdmin[v], dmn), atomicMax(&L.cdmax[v], dmx);
            }
            if (f == o) {
                atomicMin(&L.dmin[c >> 10], (unsigned long long)x[o]);
                atomicMax(&L.dmax[c >> 10], (unsigned long long)x[o]);
            }
        }
        __syncthreads();
        if (t == 0) {
            uint64_t hzl = 0, pos = 0;
            while (D - pos >= 1024) {
                const uint32_t bw = bitlen64(L.dmax[pos >> 10] - L.dmin[pos >> 10]);
                if (bw >= 64) break;
                hzl += 12 + 128ull * bw;
                pos += 1024;
            }
            hzl += 12 + 8ull * (D - pos);
            const uint64_t dict = 8 + hzl;
            uint64_t       s[kK] = {};
            for (unsigned v = 0; v < kVPR; ++v) {
                s[0] += 8192;
                s[1] += 9 + 128ull * L.bwf[v];
                s[2] += 137 + 128ull * L.bwd[v];
                s[3] += 8ull * L.runs[v] + 131 + 128ull * L.bwp[v];
                s[4] += 8ull * L.runs[v] + 2048;
                s[5] += 4096;
                s[6] += 5 + 128ull * bitlen64(L.cmax[v] - L.cmin[v]);
                s[7] += 133 + 128ull * bitlen64(L.cdmax[v] - L.cdmin[v]);
            }
            constexpr unsigned arr[kK] = {1, 3, 4, 5, 2, 1, 3, 4};
            for (unsigned k = 0; k < kK; ++k) {
                s[k] += 8 + 4ull * arr[k] * kVPR + (k >= 5 ? dict : 0);
                sizes[r * kK + k] = s[k];
            }
        }
        // reset the slots this row-group used (the next row-group of this block reuses the table)
        for (unsigned o = t; o < kRG; o += blockDim.x) {
            const uint32_t h = slot[o];
            keys[h]          = kEmpty;  // slot kHt (the kEmpty value's own slot) is padding here
            first[h]         = ~0u;
        }
        __threadfence();
        __syncthreads();
    }
}

__global__ void k_wizard_best(const uint64_t* __restrict__ sizes, uint64_t n_rg, int32_t* __restrict__ choice) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rg) return;
    uint64_t best = sizes[r * kK];
    int32_t  b    = 0;
    for (unsigned k = 1; k < kK; ++k)
        if (sizes[r * kK + k] < best) best = sizes[r * kK + k], b = (int32_t)k;
    choice[r] = b;
}

// ---------------------------------------------------------------------------------------------------------------
// Learned selector's statistics: one 256-thread block per row-group, reading the 4 sample vectors only.
// ---------------------------------------------------------------------------------------------------------------
constexpr unsigned kSHt = kSample;  // 4096 slots (+1 for the kEmpty value): never full for 4096 inputs
struct FeatLds {
    unsigned long long keys[kSHt];
    uint32_t           first[kSHt + 1];
    uint16_t           slot[kSample];
    uint32_t           bitmap[kSample / 32], base[kSample / 32];
    uint32_t           bwf[kSampleV], bwd[kSampleV], runs[kSampleV], bwp[kSampleV], eq[kSampleV], nondec[kSampleV],
        inc[kSampleV];
    unsigned long long umin[kSampleV], umax[kSampleV];
    long long          smin[kSampleV], smax[kSampleV];
    uint32_t           cmin[kSampleV], cmax[kSampleV], cdmin[kSampleV], cdmax[kSampleV];
    uint32_t           top;
    uint32_t           tmp[8];
};

__device__ __forceinline__ float log2f_(float x) { return __log2f(x); }

__global__ __launch_bounds__(256) void k_wizard_features(const uint64_t* __restrict__ values, uint64_t n_rg,
                                                         float* __restrict__ feat) {
    __shared__ FeatLds L;
    const unsigned t = threadIdx.x, lane = t % 64, wave = t / 64;
    for (uint64_t r = blockIdx.x; r < n_rg; r += gridDim.x) {
        const uint64_t* x = values + r * kRG;
        for (unsigned i = t; i < kSHt; i += blockDim.x) L.keys[i] = kEmpty, L.first[i] = ~0u;
        if (t == 0) L.first[kSHt] = ~0u, L.top = 0;
        for (unsigned i = t; i < kSample / 32; i += blockDim.x) L.bitmap[i] = 0;
        if (t < kSampleV) L.cmin[t] = ~0u, L.cmax[t] = 0, L.cdmin[t] = ~0u, L.cdmax[t] = 0;
        __syncthreads();
        {  // wave w: sample vector w
            const VecStats s = vector_stats(x + kSampleVec[wave] * 1024);
            if (lane == 0) {
                L.bwf[wave] = s.bwf, L.bwd[wave] = s.bwd, L.runs[wave] = s.runs, L.bwp[wave] = s.bwp;
                L.eq[wave] = s.eq, L.nondec[wave] = s.nondec, L.inc[wave] = s.inc;
                L.umin[wave] = s.umin, L.umax[wave] = s.umax, L.smin[wave] = s.smin, L.smax[wave] = s.smax;
            }
        }
        uint64_t xv[kSample / 256];
#pragma unroll
        for (unsigned j = 0; j < kSample / 256; ++j) {
            const unsigned p = t + 256 * j;
            xv[j]            = x[kSampleVec[p >> 10] * 1024 + (p & 1023)];
            L.slot[p]        = (uint16_t)ht_insert(L.keys, L.first, kSHt - 1, xv[j], p);
        }
        __syncthreads();
#pragma unroll
        for (unsigned j = 0; j < kSample / 256; ++j) {
            const unsigned p = t + 256 * j;
            if (L.first[L.slot[p]] == p) atomicOr(&L.bitmap[p >> 5], 1u << (p & 31));
        }
        __syncthreads();
        const uint32_t D = block_popcount_scan(L.bitmap, L.base, kSample / 32, L.tmp);
        uint32_t       fp[kSample / 256];
#pragma unroll
        for (unsigned j = 0; j < kSample / 256; ++j) {
            const unsigned p  = t + 256 * j;
            fp[j]             = L.first[L.slot[p]];
            const uint32_t c  = rank_of(L.bitmap, L.base, fp[j]);
            const uint32_t cp = (p % 32) ? rank_of(L.bitmap, L.base, L.first[L.slot[p - 1]]) : c;
            const uint32_t d  = c - cp;
            const uint32_t cmn = wave_min(c), cmx = wave_max(c), dmn = wave_min(d), dmx = wave_max(d);
            if (lane == 0) {
                const unsigned v = p >> 10;
                atomicMin(&L.cmin[v], cmn), atomicMax(&L.cmax[v], cmx);
                atomicMin(&L.cdmin[v], dmn), atomicMax(&L.cdmax[v], dmx);
            }
        }
        __syncthreads();
        // multiplicities: reuse first[] as counters
        for (unsigned i = t; i <= kSHt; i += blockDim.x) L.first[i] = 0;
        __syncthreads();
#pragma unroll
        for (unsigned j = 0; j < kSample / 256; ++j) atomicAdd(&L.first[L.slot[t + 256 * j]], 1u);
        __syncthreads();
        {
            uint32_t m = 0;
            for (unsigned i = t; i <= kSHt; i += blockDim.x) m = L.first[i] > m ? L.first[i] : m;
            m = wave_max(m);
            if (lane == 0) atomicMax(&L.top, m);
        }
        __syncthreads();
        if (t == 0) {
            float*         F  = feat + r * 64;
            const float    S  = (float)kSampleV, vv = (float)kVPR;
            uint32_t       sbwf = 0, sbwd = 0, sruns = 0, sbwp = 0, sbwc = 0, sbwcd = 0, eq = 0, nondec = 0, inc = 0;
            uint32_t       mxbwf = 0, mnbwf = 64, mxbwd = 0, mnbwd = 64, mxr = 0, mnr = 1024;
            uint64_t       umax = 0, umin = kEmpty;
            int64_t        smax = INT64_MIN, smin = INT64_MAX;
            uint32_t       bwc[kSampleV], bwcd[kSampleV];
            for (unsigned i = 0; i < kSampleV; ++i) {
                bwc[i]  = bitlen64(L.cmax[i] - L.cmin[i]);
                bwcd[i] = bitlen64(L.cdmax[i] - L.cdmin[i]);
                sbwf += L.bwf[i], sbwd += L.bwd[i], sruns += L.runs[i], sbwp += L.bwp[i];
                sbwc += bwc[i], sbwcd += bwcd[i];
                eq += L.eq[i], nondec += L.nondec[i], inc += L.inc[i];
                mxbwf = max(mxbwf, L.bwf[i]), mnbwf = min(mnbwf, L.bwf[i]);
                mxbwd = max(mxbwd, L.bwd[i]), mnbwd = min(mnbwd, L.bwd[i]);
                mxr = max(mxr, L.runs[i]), mnr = min(mnr, L.runs[i]);
                umax = L.umax[i] > umax ? L.umax[i] : umax;
                umin = L.umin[i] < umin ? L.umin[i] : umin;
                smax = L.smax[i] > smax ? L.smax[i] : smax;
                smin = L.smin[i] < smin ? L.smin[i] : smin;
            }
            const float mbwf = (float)sbwf / S, mbwd = (float)sbwd / S, mruns = (float)sruns / S,
                        mbwp = (float)sbwp / S, mbwc = (float)sbwc / S, mbwcd = (float)sbwcd / S;
            const float fD = (float)D;
            float       est[kK];
            est[0] = 8.f + 4.f * vv + 8192.f * vv;
            est[1] = 8.f + 12.f * vv + vv * (9.f + 128.f * mbwf);
            est[2] = 8.f + 16.f * vv + vv * (137.f + 128.f * mbwd);
            est[3] = 8.f + 20.f * vv + vv * (8.f * mruns + 131.f + 128.f * mbwp);
            est[4] = 8.f + 8.f * vv + vv * (8.f * mruns + 2048.f);
            est[5] = 28.f + 8.f * fD + 4.f * vv + 4096.f * vv;
            est[6] = 28.f + 8.f * fD + 12.f * vv + vv * (5.f + 128.f * mbwc);
            est[7] = 28.f + 8.f * fD + 16.f * vv + vv * (133.f + 128.f * mbwcd);
            for (unsigned k = 0; k < kK; ++k) F[k] = est[k] / (float)(kRG * 8), F[8 + k] = log2f_(est[k]) / 24.f;
            F[16] = mbwf / 64.f, F[17] = (float)mxbwf / 64.f, F[18] = (float)mnbwf / 64.f;
            F[19] = mbwd / 64.f, F[20] = (float)mxbwd / 64.f, F[21] = (float)mnbwd / 64.f;
            F[22] = mruns / 1024.f, F[23] = (float)mxr / 1024.f, F[24] = (float)mnr / 1024.f;
            F[25] = log2f_(mruns) / 10.f;
            F[26] = mbwp;
            F[27] = fD / (S * 1024.f);
            F[28] = log2f_(fD) / 12.f;
            F[29] = (float)bitlen64(D - 1) / 16.f;
            F[30] = (float)L.top / (S * 1024.f);
            F[31] = (float)eq / (S * 1023.f);
            F[32] = (float)nondec / (S * 1023.f);
            F[33] = (float)inc / (S * 1023.f);
            F[34] = (float)bitlen64(umax) / 64.f;
            F[35] = (float)bitlen64(umin) / 64.f;
            F[36] = (float)bitlen64((uint64_t)smax - (uint64_t)smin) / 64.f;
            F[37] = mbwc / 32.f;
            F[38] = mbwcd / 32.f;
            for (unsigned i = 0; i < kSampleV; ++i) {
                F[39 + 6 * i] = (float)L.bwf[i] / 64.f, F[40 + 6 * i] = (float)L.bwd[i] / 64.f;
                F[41 + 6 * i] = (float)L.runs[i] / 1024.f, F[42 + 6 * i] = (float)L.bwp[i];
                F[43 + 6 * i] = (float)bwc[i] / 32.f, F[44 + 6 * i] = (float)bwcd[i] / 32.f;
            }
            F[63] = 1.f;
        }
        __syncthreads();
    }
}

// per-device workspace of the exhaustive sizer
struct SizerWs {
    uint64_t* tables = nullptr;
    uint32_t* firsts = nullptr;
    uint32_t* slots  = nullptr;
    unsigned  blocks = 0;
};
std::mutex           g_ws_mu;
std::vector<SizerWs> g_ws;

int sizer_ws(SizerWs*& out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_ws_mu);
    if ((int)g_ws.size() <= dev) g_ws.resize(dev + 1);
    SizerWs& w = g_ws[dev];
    if (!w.tables) {
        hipDeviceProp_t prop;
        if ((e = hipGetDeviceProperties(&prop, dev)) != hipSuccess) return hip_fail(e, "device properties");
        w.blocks = (unsigned)prop.multiProcessorCount;
        const size_t tb = (size_t)w.blocks * (kHt + 64) * 8, fb = (size_t)w.blocks * (2 * kHt + 64) * 4,
                     sb = (size_t)w.blocks * kRG * 4;
        if ((e = hipMalloc(&w.tables, tb)) != hipSuccess) return hip_fail(e, "hipMalloc");
        if ((e = hipMalloc(&w.firsts, fb)) != hipSuccess) return hip_fail(e, "hipMalloc");
        if ((e = hipMalloc(&w.slots, sb)) != hipSuccess) return hip_fail(e, "hipMalloc");
        // empty keys and first positions; the kernels restore this state after every row-group
        if ((e = hipMemset(w.tables, 0xFF, tb)) != hipSuccess) return hip_fail(e, "hipMemset");
        if ((e = hipMemset(w.firsts, 0xFF, fb)) != hipSuccess) return hip_fail(e, "hipMemset");
        if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(e, "sync");
    }
    out = &w;
    return FLS_OK;
}

}  // namespace
}  // namespace fls

using namespace fls;

extern "C" {

uint32_t fls_wizard_n_candidates(void) { return kK; }
uint8_t  fls_wizard_candidate_expr(uint32_t k) { return k < kK ? kCandExpr[k] : 255; }

int fls_wizard_sizes(const int64_t* dev_values, uint64_t n_rowgroups, uint64_t* dev_sizes, void* stream) {
    if (!n_rowgroups) return FLS_OK;
    if (!dev_values || !dev_sizes) return fail(FLS_E_ARG, "null argument");
    if ((uintptr_t)dev_values & 15) return fail(FLS_E_ARG, "values must be 16-byte aligned");
    SizerWs* w  = nullptr;
    int      rc = sizer_ws(w);
    if (rc) return rc;
    const unsigned blocks = (unsigned)std::min<uint64_t>(n_rowgroups, w->blocks);
    hipLaunchKernelGGL(k_wizard_sizes, dim3(blocks), dim3(1024), 0, (hipStream_t)stream, (const uint64_t*)dev_values,
                       n_rowgroups, w->tables, w->firsts, w->slots, dev_sizes);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "wizard sizes launch");
}

int fls_wizard_best(const uint64_t* dev_sizes, uint64_t n_rowgroups, int32_t* dev_choice, void* stream) {
    if (!n_rowgroups) return FLS_OK;
    if (!dev_sizes || !dev_choice) return fail(FLS_E_ARG, "null argument");
    hipLaunchKernelGGL(k_wizard_best, dim3((unsigned)((n_rowgroups + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       dev_sizes, n_rowgroups, dev_choice);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "wizard best launch");
}

int fls_wizard_features(const int64_t* dev_values, uint64_t n_rowgroups, float* dev_features, void* stream) {
    if (!n_rowgroups) return FLS_OK;
    if (!dev_values || !dev_features) return fail(FLS_E_ARG, "null argument");
    if ((uintptr_t)dev_values & 15) return fail(FLS_E_ARG, "values must be 16-byte aligned");
    const unsigned blocks = (unsigned)std::min<uint64_t>(n_rowgroups, 65535u);
    hipLaunchKernelGGL(k_wizard_features, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       (const uint64_t*)dev_values, n_rowgroups, dev_features);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "wizard features launch");
}

}  // extern "C"
