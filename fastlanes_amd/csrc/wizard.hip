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

constexpr unsigned kRG = 65536, kVPR = 64, kK = 16;
constexpr uint64_t kEmpty = ~0ull;
constexpr unsigned kHtBits = 17, kHt = 1u << kHtBits;      // exhaustive table: 131072 slots (load <= 0.5)
constexpr unsigned kSampleV = 4, kSample = kSampleV * 1024;  // learned selector's sample
constexpr unsigned kSampleVec[kSampleV] = {0, 21, 42, 63};
constexpr uint8_t kCandExpr[kK] = {0, 1, 11, 6, 7, 16, 17, 22, 2, 3, 4, 5, 27, 12, 13, 14};
constexpr unsigned kExcLimit = 100;  // CFG::CMPR::EXC_LIMIT_C (include/fls/cfg/cfg.hpp:67)
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

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sorts the 1024 values of buf (the wave's LDS buffer) ascending with a bitonic network held in registers: lane l
// keeps elements 16l .. 16l+15, so the stages over distances 1-8 (34 of 55) are compare-exchanges inside a lane and
// the stages over distances 16-512 exchange with lane l ^ (distance / 16) through shuffles; LDS is read once and
// written once.
__device__ __forceinline__ void ce_regs(uint64_t& a, uint64_t& b, bool desc) {
    const bool     sw = (a > b) != desc;
    const uint64_t x = sw ? b : a, y = sw ? a : b;
    a = x, b = y;
}
__device__ void wave_sort1024(uint64_t* buf) {
    const unsigned lane = threadIdx.x % 64;
    uint64_t       v[16];
    wave_lds_sync();
#pragma unroll
    for (unsigned r = 0; r < 16; ++r) v[r] = buf[64 * r + lane];  // any starting order will do
#pragma unroll
    for (unsigned k = 2; k <= 1024; k <<= 1) {
        // descending blocks of size k: element i = 16 lane + r with (i & k) != 0
#pragma unroll
        for (unsigned j = k >> 1; j > 0; j >>= 1) {
            if (j >= 16) {
                const unsigned d       = j / 16;
                const bool     desc    = k < 1024 && (((16 * lane) & k) != 0);
                const bool     takemin = ((lane & d) == 0) != desc;  // the lower index of an ascending pair
#pragma unroll
                for (unsigned r = 0; r < 16; ++r) {
                    const uint64_t p = __shfl_xor((unsigned long long)v[r], (int)d, 64);
                    v[r]             = ((v[r] < p) == takemin) ? v[r] : p;
                }
            } else {
#pragma unroll
                for (unsigned r = 0; r < 16; ++r) {
                    if (r & j) continue;
                    const bool desc = k < 16 ? (r & k) != 0 : (k < 1024 && ((16 * lane) & k) != 0);
                    ce_regs(v[r], v[r | j], desc);
                }
            }
        }
    }
#pragma unroll
    for (unsigned r = 0; r < 16; ++r) buf[16 * lane + r] = v[r];
    wave_lds_sync();
}

// Original position of transposed index k (the 1024-value transpose of every FastLanes type: k = 128r + 16b + c
// holds value r + 8 ORDER[b] + 64c).
__device__ __forceinline__ unsigned orig_of(unsigned k) {
    const unsigned b = (k >> 4) & 7;
    return (k >> 7) + 8 * (((b & 1) << 2) | (b & 2) | (b >> 2)) + 64 * (k & 15);
}

// Patch plan of one vector (find_best_params, src/cor/prm/ffor/ffor_analyze.cpp:41-75, as the host encoder's
// plan_patch): over bit widths bw < 64 and window bases s[i] (first index of each distinct value among the 101
// smallest), the window [s[i], s[i] + 2^bw - 1] (bw = 0: only base 0, window {0}) leaving at most 100 values
// outside; the plan minimises 128 bw + 8 (values outside), ties to the smaller i, then the larger bw; it exists only
// below 128 * 64 bytes (else bw = 64, base 0, no exceptions). s: the vector sorted ascending (LDS). A window wider
// than the vector's unsigned FFOR width bwu is never better than bw = bwu at i = 0, so bw runs up to min(bwu, 63) --
// and up to 1 for a constant nonzero vector, which no zero-width window holds.
struct PatchPlan {
    uint32_t bw, exc;
    uint64_t base;
};
__device__ PatchPlan patch_plan(const uint64_t* s, uint32_t bwu) {
    const unsigned lane = threadIdx.x % 64, bw_hi = bwu < 1 ? 1 : bwu < 63 ? bwu : 63;
    uint64_t       key  = ~0ull;  // (size << 32) | (i << 8) | (63 - bw): the smallest key is the plan
    for (unsigned h = 0; h < 2; ++h) {
        const unsigned i = lane + 64 * h;
        if (i > kExcLimit) continue;
        const uint64_t si = s[i];
        if (i && si == s[i - 1]) continue;
        // the window must hold s[i .. 923 + i]: its width is at least bitlen(s[923 + i] - s[i])
        const uint32_t bw_lo = bitlen64(s[1023 - kExcLimit + i] - si);
        for (unsigned bw = bw_lo; bw <= bw_hi; ++bw) {
            uint64_t upper;
            if (bw == 0) {
                if (si != 0) continue;  // a zero-width window holds only zeros
                upper = 0;
            } else {
                const uint64_t range = (1ull << bw) - 1;
                upper = si > ~0ull - range ? ~0ull : si + range;
            }
            // first index past the window, searched above 923 + i (every smaller index is inside it)
            unsigned lo = 1023 - kExcLimit + i, hi = 1024;
            while (hi - lo > 1) {
                const unsigned mid = (lo + hi) / 2;
                if (s[mid] <= upper) lo = mid;
                else hi = mid;
            }
            const uint32_t exc  = 1024 - (hi - i);
            const uint64_t size = 128ull * bw + 8ull * exc;
            const uint64_t kk   = (size << 32) | ((uint64_t)i << 8) | (63 - bw);
            key = kk < key ? kk : key;
        }
    }
    key = wave_min(key);
    PatchPlan pp {64, 0, 0};
    if ((key >> 32) < 128ull * 64) {
        const unsigned i = (unsigned)(key >> 8) & 0xffffff;
        pp.bw   = 63 - (uint32_t)(key & 0xff);
        pp.base = s[i];
        pp.exc  = (uint32_t)(((key >> 32) - 128ull * pp.bw) / 8);
    }
    return pp;
}
__device__ __forceinline__ bool is_exception(uint64_t v, const PatchPlan& pp) {
    if (pp.bw >= 64) return false;
    if (pp.bw == 0) return v != 0;
    const uint64_t range = (1ull << pp.bw) - 1;
    const uint64_t upper = pp.base > ~0ull - range ? ~0ull : pp.base + range;
    return v < pp.base || v > upper;
}

// Statistics of one 1024-value vector, computed by one wave (lane l owns values 16l .. 16l+15).
// bwf / bwd: FFOR widths of the values and of their lane differences as ANALYZE computes them for int64 columns
// (min and max signed, ffor_analyze<i64_pt>); bwu / bwdu: their unsigned widths (bound the patch searches).
struct VecStats {
    uint32_t bwf, bwd, bwu, bwdu, runs, bwp, eq, nondec, inc;
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
    uint64_t mn = kEmpty, mx = 0, dun = kEmpty, dux = 0;
    int64_t  smn = INT64_MAX, smx = INT64_MIN, dmn = INT64_MAX, dmx = INT64_MIN;
    uint32_t runs = 0, inner = 0, eq = 0, nondec = 0, inc = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const unsigned o  = 16 * lane + i;
        const uint64_t pv = i ? v[i - 1] : prev;
        mn = v[i] < mn ? v[i] : mn;
        mx = v[i] > mx ? v[i] : mx;
        smn = (int64_t)v[i] < smn ? (int64_t)v[i] : smn;
        smx = (int64_t)v[i] > smx ? (int64_t)v[i] : smx;
        const int64_t  d = (o % 64) ? (int64_t)(v[i] - pv) : 0;
        dun = (uint64_t)d < dun ? (uint64_t)d : dun;
        dux = (uint64_t)d > dux ? (uint64_t)d : dux;
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
    s.bwf  = bitlen64((uint64_t)s.smax - (uint64_t)s.smin);
    s.bwu  = bitlen64(s.umax - s.umin);
    s.bwd  = bitlen64((uint64_t)wave_max(dmx) - (uint64_t)wave_min(dmn));
    s.bwdu = bitlen64(wave_max(dux) - wave_min(dun));
    s.runs = wave_sum(runs), s.bwp = wave_sum(inner) != 0;
    s.eq = wave_sum(eq), s.nondec = wave_sum(nondec), s.inc = wave_sum(inc);
    return s;
}

// Atomics on a block's own region of global memory (the sizer's hash table): only the block's threads touch it,
// so workgroup scope is enough, and the operations stay in the XCD's L2 instead of the device-coherent path.
__device__ __forceinline__ uint32_t wg_cas(uint32_t* p, uint32_t expected, uint32_t desired) {
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    return expected;  // the value found (equal to the expected one when the exchange happened)
}
__device__ __forceinline__ void wg_min(uint32_t* p, uint32_t v) {
    __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t l2_load(const uint32_t* p) {  // reads the L2, where the exchanges happen
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The sizer's open-addressing table keeps, per slot, the smallest position of its value seen so far (kNoPos: empty);
// keys compare through the row-group's own values (x[position], resident in the L2 while the row-group is sized),
// so a slot is 4 bytes and the block's table 512 KiB. A claimed slot only ever holds positions of one value.
constexpr uint32_t kNoPos = ~0u;
// Probes from slot h, whose content was read as cur, for value v at position o; returns v's slot.
__device__ __forceinline__ uint32_t ht_insert(uint32_t* tab, const uint64_t* __restrict__ x, uint32_t mask, uint64_t v,
                                              uint32_t o, uint32_t h, uint32_t cur) {
    for (uint32_t p = 0; p <= mask; ++p) {
        if (cur == kNoPos) {
            cur = wg_cas(&tab[h], kNoPos, o);
            if (cur == kNoPos) return h;  // claimed
        }
        if (x[cur] == v) {
            if (o < cur) wg_min(&tab[h], o);
            return h;
        }
        h   = (h + 1) & mask;
        cur = l2_load(&tab[h]);
    }
    return 0;  // unreachable: the table holds more slots than distinct inputs
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
// Only the block's own threads touch its table, slots and first positions, so the barriers' workgroup-scope
// ordering is all the phases need (a device-scope fence would write back and invalidate the XCD's whole L2).
// ---------------------------------------------------------------------------------------------------------------
// DELTA patches (delta_exp.hpp:56-131: RSUM, then PA_ANALYZE over the lane differences): the patch plan of one
// vector's differences (the value minus its predecessor inside each 64-value lane, 0 at a lane start), sorted in the
// wave's buffer sb, and its exceptions in transposed order with the LL chain's hop links, as for the values.
__device__ PatchPlan delta_patch(const uint64_t* __restrict__ xv, uint64_t* sb, uint32_t bwdu, uint32_t& n,
                                 uint32_t& extra) {
    const unsigned lane = threadIdx.x % 64;
    auto diff = [&](unsigned o) { return (o % 64) ? xv[o] - xv[o - 1] : 0ull; };
    wave_lds_sync();  // the previous plan has read the sort buffer
#pragma unroll
    for (unsigned i = 0; i < 16; ++i) sb[64 * i + lane] = diff(64 * i + lane);
    wave_sort1024(sb);
    const PatchPlan pd  = patch_plan(sb, bwdu);
    const unsigned  hop = 1u << (pd.bw < 11 ? pd.bw : 11);
    int             last = -1;
    n = 0, extra = 0;
    uint64_t dq[16];  // the 16 differences this lane tests, loaded before the ballot chain
#pragma unroll
    for (unsigned q = 0; q < 16; ++q) dq[q] = diff(orig_of(64 * q + lane));
#pragma unroll
    for (unsigned q = 0; q < 16; ++q) {
        const unsigned k  = 64 * q + lane;
        const bool     ex = is_exception(dq[q], pd);
        const uint64_t m  = __ballot(ex);
        if (ex) {
            const uint64_t below = m & ((1ull << lane) - 1);
            const int      prev  = below ? (int)(64 * q + 63 - __clzll((long long)below)) : last;
            if (prev >= 0) extra += (k - (unsigned)prev - 1) / hop;
        }
        if (m) last = (int)(64 * q + 63 - __clzll((long long)m));
        n += (uint32_t)__popcll(m);
    }
    extra = wave_sum(extra);
    return pd;
}

struct SizerLds {
    uint32_t bitmap[kRG / 32];
    uint32_t base[kRG / 32];
    uint32_t bwf[kVPR], bwd[kVPR], runs[kVPR], bwp[kVPR];
    uint32_t cmin[kVPR], cmax[kVPR], cdmin[kVPR], cdmax[kVPR];
    unsigned long long dmin[kVPR], dmax[kVPR];  // dictionary chunk min / max (chunks of 1024 codes)
    uint32_t tmp[32];
    // patched FFOR: per vector bit width, base, exceptions and LL chain length; the page's exception values in
    // chunks of 1024 (S_PATCH_REC's HZL chain)
    uint32_t           pbw[kVPR], pexc[kVPR], pchain[kVPR], excpref[kVPR + 1];
    // DELTA patches: the same plan over each vector's lane differences
    uint32_t           dbw[kVPR], dexc[kVPR], dchain[kVPR];
    unsigned long long pbase[kVPR], excmin[8], excmax[8];
    // DICT_RLE: the page's run codes in chunks of 1024 (its HZL chain)
    uint32_t rcmin[kVPR], rcmax[kVPR], runpref[kVPR + 1];
    uint64_t sortbuf[16][1024];  // one vector per wave, sorted for the patch plan
};

__global__ __launch_bounds__(1024) void k_wizard_sizes(const uint64_t* __restrict__ values, uint64_t n_rg,
                                                       uint32_t* __restrict__ firsts_all, uint32_t* __restrict__ slots_all,
                                                       uint64_t* __restrict__ sizes) {
    // dynamic LDS (the launch passes sizeof(SizerLds), about 150 KB)
    extern __shared__ __attribute__((aligned(16))) uint8_t wiz_lds[];
    SizerLds&      L = *reinterpret_cast<SizerLds*>(wiz_lds);
    const unsigned t = threadIdx.x, lane = t % 64, wave = t / 64;
    uint32_t* first  = firsts_all + (uint64_t)blockIdx.x * (2 * kHt + 64);  // the table: first position per slot
    uint32_t* fpos   = first + kHt + 32;                                   // first occurrence per position (kRG)
    uint32_t* slot   = slots_all + (uint64_t)blockIdx.x * kRG;
    for (uint64_t r = blockIdx.x; r < n_rg; r += gridDim.x) {
        const uint64_t* x = values + r * kRG;
        for (unsigned i = t; i < kRG / 32; i += blockDim.x) L.bitmap[i] = 0;
        if (t < kVPR) {
            L.cmin[t] = ~0u, L.cmax[t] = 0, L.cdmin[t] = ~0u, L.cdmax[t] = 0;
            L.dmin[t] = kEmpty, L.dmax[t] = 0;
            L.rcmin[t] = ~0u, L.rcmax[t] = 0;
        }
        if (t < 8) L.excmin[t] = kEmpty, L.excmax[t] = 0;
        // per-vector statistics and patch plans (wave w takes vectors w, w + 16, ...)
        for (unsigned v = wave; v < kVPR; v += blockDim.x / 64) {
            const VecStats s = vector_stats(x + v * 1024);
            uint64_t*      sb = L.sortbuf[wave];
#pragma unroll
            for (unsigned i = 0; i < 16; ++i) sb[64 * i + lane] = x[v * 1024 + 64 * i + lane];
            wave_sort1024(sb);
            const PatchPlan pp = patch_plan(sb, s.bwu);
            // exceptions in transposed order (the order the encoder stores them) and the LL chain's hop entries:
            // a gap g between consecutive exceptions adds floor((g - 1) / hop) links, hop = 2^min(bw, 11)
            const unsigned hop = 1u << (pp.bw < 11 ? pp.bw : 11);
            uint32_t       n = 0, extra = 0;
            int            last = -1;  // last exception index seen by the wave so far
            uint64_t       xq[16];     // the 16 values this lane tests, loaded before the ballot chain
#pragma unroll
            for (unsigned q = 0; q < 16; ++q) xq[q] = x[v * 1024 + orig_of(64 * q + lane)];
#pragma unroll
            for (unsigned q = 0; q < 16; ++q) {
                const unsigned k  = 64 * q + lane;
                const bool     ex = is_exception(xq[q], pp);
                const uint64_t m  = __ballot(ex);
                if (ex) {
                    const uint64_t below = m & ((1ull << lane) - 1);
                    const int prev = below ? (int)(64 * q + 63 - __clzll((long long)below)) : last;
                    if (prev >= 0) extra += (k - (unsigned)prev - 1) / hop;
                }
                if (m) last = (int)(64 * q + 63 - __clzll((long long)m));
                n += (uint32_t)__popcll(m);
            }
            extra = wave_sum(extra);
            if (lane == 0) {
                L.bwf[v] = s.bwf, L.bwd[v] = s.bwd, L.runs[v] = s.runs, L.bwp[v] = s.bwp;
                L.pbw[v] = pp.bw, L.pbase[v] = pp.base, L.pexc[v] = n, L.pchain[v] = n + extra;
            }
            uint32_t dn = 0, dextra = 0;
            const PatchPlan pd = delta_patch(x + v * 1024, sb, s.bwdu, dn, dextra);
            if (lane == 0) L.dbw[v] = pd.bw, L.dexc[v] = dn, L.dchain[v] = dn + dextra;
        }
        // dictionary: insert every value, remember its slot. A thread takes kBatch positions at a time: their loads
        // and first probes are issued together, so kBatch atomics are in flight per thread instead of one.
        constexpr unsigned kBatch = 4;
        const unsigned     stride = blockDim.x;
        for (unsigned o0 = t; o0 < kRG; o0 += kBatch * stride) {
            uint64_t v[kBatch];
            uint32_t h[kBatch], cur[kBatch], hs[kBatch];
            bool     lead[kBatch], done[kBatch];
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k) v[k] = x[o0 + k * stride];
            // a run of equal values at consecutive positions of the wave: only its first lane inserts (the others'
            // positions are no first occurrences) and the others take its slot -- many lanes exchanging or
            // min-ing one slot would serialize on that address
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k) {
                const uint64_t pv = __shfl_up(v[k], 1, 64);
                lead[k]           = lane == 0 || pv != v[k];
                h[k]              = mix64(v[k]) & (kHt - 1);
                hs[k]             = h[k];
            }
            // the home slots' contents, then the exchanges into the empty ones, then the value comparisons: the
            // batch's loads and atomics of each step are in flight together
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k) cur[k] = lead[k] ? l2_load(&first[h[k]]) : 0u;
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k) {
                done[k] = !lead[k];
                if (!done[k] && cur[k] == kNoPos) {
                    cur[k]  = wg_cas(&first[h[k]], kNoPos, o0 + k * stride);
                    done[k] = cur[k] == kNoPos;
                }
            }
            uint64_t xv[kBatch];
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k) xv[k] = done[k] ? 0ull : x[cur[k]];
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k) {
                const unsigned o = o0 + k * stride;
                if (done[k]) continue;
                if (xv[k] == v[k]) {
                    if (o < cur[k]) wg_min(&first[h[k]], o);
                } else {  // the home slot holds another value: probe on
                    hs[k] = ht_insert(first, x, kHt - 1, v[k], o, (h[k] + 1) & (kHt - 1),
                                      l2_load(&first[(h[k] + 1) & (kHt - 1)]));
                }
            }
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k) {
                const uint64_t lm  = __ballot(lead[k]);  // lane 0 always leads
                const unsigned src = 63u - (unsigned)__clzll((long long)(lm & ((2ull << lane) - 1)));
                slot[o0 + k * stride] = (uint32_t)__shfl((int)hs[k], (int)src, 64);
            }
        }
        __syncthreads();
        for (unsigned o0 = t; o0 < kRG; o0 += kBatch * stride) {
            uint32_t sl[kBatch], f[kBatch];
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k) sl[k] = slot[o0 + k * stride];
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k)
                f[k] = l2_load(&first[sl[k]]);
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k) {
                const unsigned o = o0 + k * stride;
                fpos[o]          = f[k];
                if (f[k] == o) atomicOr(&L.bitmap[o >> 5], 1u << (o & 31));
            }
        }
        __syncthreads();
        const uint32_t D = block_popcount_scan(L.bitmap, L.base, kRG / 32, L.tmp);
        // codes: per vector min/max and 32-segment delta min/max; dictionary chunk min/max
        for (unsigned o0 = t; o0 < kRG; o0 += kBatch * stride) {
          uint32_t fb[kBatch], fpb[kBatch];
#pragma unroll
          for (unsigned k = 0; k < kBatch; ++k) {
              const unsigned o = o0 + k * stride;
              fb[k]            = fpos[o];
              fpb[k]           = (o % 32) ? fpos[o - 1] : 0u;
          }
#pragma unroll
          for (unsigned k = 0; k < kBatch; ++k) {
            const unsigned o  = o0 + k * stride;
            const uint32_t f  = fb[k];
            const uint32_t c  = rank_of(L.bitmap, L.base, f);
            const uint32_t cp = (o % 32) ? rank_of(L.bitmap, L.base, fpb[k]) : c;
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
        }
        __syncthreads();
        if (t == 0) {
            uint32_t a = 0, b = 0;
            for (unsigned v = 0; v < kVPR; ++v) L.runpref[v] = a, a += L.runs[v], L.excpref[v] = b, b += L.pexc[v];
            L.runpref[kVPR] = a, L.excpref[kVPR] = b;
        }
        __syncthreads();
        // DICT_RLE run codes and S_PATCH_REC exception values, each as the page's chain in chunks of 1024
        for (unsigned v = wave; v < kVPR; v += blockDim.x / 64) {
            const uint64_t* xv = x + v * 1024;
            PatchPlan       pp {L.pbw[v], 0, L.pbase[v]};
            uint32_t        nr = L.runpref[v], ne = L.excpref[v];
            for (unsigned q = 0; q < 16; ++q) {
                const unsigned o   = 64 * q + lane;
                const uint64_t val = xv[o];
                const bool     chg = o == 0 || val != xv[o - 1];
                const uint64_t mr  = __ballot(chg);
                if (chg) {
                    const uint32_t c     = rank_of(L.bitmap, L.base, fpos[v * 1024 + o]);
                    const uint32_t chunk = (nr + (uint32_t)__popcll(mr & ((1ull << lane) - 1))) >> 10;
                    atomicMin(&L.rcmin[chunk], c), atomicMax(&L.rcmax[chunk], c);
                }
                nr += (uint32_t)__popcll(mr);
                const uint64_t ev = xv[orig_of(o)];
                const bool     ex = is_exception(ev, pp);
                const uint64_t me = __ballot(ex);
                if (ex) {
                    const uint32_t chunk = (ne + (uint32_t)__popcll(me & ((1ull << lane) - 1))) >> 10;
                    atomicMin(&L.excmin[chunk], (unsigned long long)ev), atomicMax(&L.excmax[chunk], (unsigned long long)ev);
                }
                ne += (uint32_t)__popcll(me);
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
            // HZL chains of the page's run codes (u32) and exception values (u64): 12 + 128 bw per full chunk of 1024
            // while bw is below the type width, then 12 + the raw rest
            uint64_t       rchain = 0, echain = 0;
            {
                const uint32_t R = L.runpref[kVPR];
                uint32_t       p = 0;
                for (; R - p >= 1024; p += 1024) {
                    const uint32_t bw = bitlen64(L.rcmax[p >> 10] - L.rcmin[p >> 10]);
                    if (bw >= 32) break;
                    rchain += 12 + 128ull * bw;
                }
                rchain += 12 + 4ull * (R - p);
                const uint32_t E = L.excpref[kVPR];
                p = 0;
                for (; E - p >= 1024; p += 1024) {
                    const uint32_t bw = bitlen64(L.excmax[p >> 10] - L.excmin[p >> 10]);
                    if (bw >= 64) break;
                    echain += 12 + 128ull * bw;
                }
                echain += 12 + 8ull * (E - p);
            }
            uint64_t s[kK] = {};
            for (unsigned v = 0; v < kVPR; ++v) {
                const uint64_t pk = 9 + 128ull * L.pbw[v], n = L.pexc[v];
                s[0] += 8192;
                s[1] += 9 + 128ull * L.bwf[v];
                s[2] += 137 + 128ull * L.bwd[v];
                s[3] += 8ull * L.runs[v] + 131 + 128ull * L.bwp[v];
                s[4] += 8ull * L.runs[v] + 2048;
                s[5] += 4096;
                s[6] += 5 + 128ull * bitlen64(L.cmax[v] - L.cmin[v]);
                s[7] += 133 + 128ull * bitlen64(L.cdmax[v] - L.cdmin[v]);
                s[8] += 8 * n + 2 + 2 * n + pk;             // S_PATCH [exceptions][count, positions][bw][base][packed]
                s[9] += 2 + 2 * n + pk;                     // S_PATCH_REC: exceptions in the page's chain
                s[10] += 8 * n + 128 + pk;                  // B_PATCH [exceptions][bitmap]...
                s[11] += 8ull * L.pchain[v] + 4 + pk;       // LL_PATCH [chain values][count, first]...
                s[12] += 131 + 128ull * L.bwp[v];           // DICT_RLE [run codes: chain][lane bases][bw][base][packed]
                const uint64_t pkd = 9 + 128ull * L.dbw[v], dn = L.dexc[v];
                s[13] += 128 + 8 * dn + 2 + 2 * dn + pkd;   // DELTA S_PATCH [lane bases][exceptions][count, positions]...
                s[14] += 128 + 8 * dn + 128 + pkd;          // DELTA B_PATCH [lane bases][exceptions][bitmap]...
                s[15] += 128 + 8ull * L.dchain[v] + 4 + pkd; // DELTA LL_PATCH [lane bases][chain values][count, first]...
            }
            s[9] += echain;
            s[12] += rchain;
            constexpr unsigned arr[kK]  = {1, 3, 4, 5, 2, 1, 3, 4, 5, 5, 5, 5, 5, 6, 6, 6};
            constexpr bool     dic[kK]  = {false, false, false, false, false, true, true, true,
                                           false, false, false, false, true, false, false, false};
            for (unsigned k = 0; k < kK; ++k) {
                s[k] += 8 + 4ull * arr[k] * kVPR + (dic[k] ? dict : 0);
                sizes[r * kK + k] = s[k];
            }
        }
        // reset the slots this row-group used (the next row-group of this block reuses the table)
        for (unsigned o0 = t; o0 < kRG; o0 += kBatch * stride) {
            uint32_t sl[kBatch];
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k) sl[k] = slot[o0 + k * stride];
#pragma unroll
            for (unsigned k = 0; k < kBatch; ++k) {
                if (fpos[o0 + k * stride] != o0 + k * stride) continue;  // each slot once: at its first occurrence
                first[sl[k]] = kNoPos;
            }
        }
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
constexpr unsigned kFHt = 2 * kSample;  // feature table: 8192 slots of a first position (load <= 0.5)
struct FeatLds {
    union {  // the sample vectors' sort buffers, then the sample values in position order
        uint64_t sortbuf[kSampleV][1024];
        uint64_t xs[kSample];
    } u;
    uint32_t           tab[kFHt];  // first position of each distinct value (then its multiplicity)
    uint16_t           slot[kSample];
    uint32_t           bitmap[kSample / 32], base[kSample / 32];
    uint32_t           bwf[kSampleV], bwd[kSampleV], runs[kSampleV], bwp[kSampleV], eq[kSampleV], nondec[kSampleV],
        inc[kSampleV];
    unsigned long long umin[kSampleV], umax[kSampleV];
    long long          smin[kSampleV], smax[kSampleV];
    uint32_t           cmin[kSampleV], cmax[kSampleV], cdmin[kSampleV], cdmax[kSampleV];
    uint32_t           top;
    uint32_t           tmp[8];
    uint32_t           pbw[kSampleV], pexc[kSampleV];    // patch plan of each sample vector
    uint32_t           dpbw[kSampleV], dpexc[kSampleV];  // and of its lane differences
};

// Inserts sample position p (value xs[p]) into the position table: a slot holds the smallest position of one value
// (values compared through xs). Returns the slot.
__device__ __forceinline__ uint32_t pos_insert(uint32_t* tab, const uint64_t* xs, uint32_t p) {
    const uint64_t v = xs[p];
    uint32_t       h = mix64(v) & (kFHt - 1);
    for (;;) {
        uint32_t cur = tab[h];
        if (cur == ~0u) {
            cur = atomicCAS(&tab[h], ~0u, p);
            if (cur == ~0u) return h;
        }
        if (xs[cur] == v) {
            atomicMin(&tab[h], p);
            return h;
        }
        h = (h + 1) & (kFHt - 1);  // never full: 4096 values, 8192 slots
    }
}

__device__ __forceinline__ float log2f_(float x) { return __log2f(x); }

__global__ __launch_bounds__(256) void k_wizard_features(const uint64_t* __restrict__ values, uint64_t n_rg,
                                                         float* __restrict__ feat) {
    extern __shared__ __attribute__((aligned(16))) uint8_t wiz_lds[];  // sizeof(FeatLds), about 74 KB
    FeatLds&       L = *reinterpret_cast<FeatLds*>(wiz_lds);
    const unsigned t = threadIdx.x, lane = t % 64, wave = t / 64;
    for (uint64_t r = blockIdx.x; r < n_rg; r += gridDim.x) {
        const uint64_t* x = values + r * kRG;
        if (t == 0) L.top = 0;
        for (unsigned i = t; i < kSample / 32; i += blockDim.x) L.bitmap[i] = 0;
        if (t < kSampleV) L.cmin[t] = ~0u, L.cmax[t] = 0, L.cdmin[t] = ~0u, L.cdmax[t] = 0;
        __syncthreads();
        {  // wave w: sample vector w
            const VecStats s  = vector_stats(x + kSampleVec[wave] * 1024);
            uint64_t*      sb = L.u.sortbuf[wave];
#pragma unroll
            for (unsigned i = 0; i < 16; ++i) sb[64 * i + lane] = x[kSampleVec[wave] * 1024 + 64 * i + lane];
            wave_sort1024(sb);
            const PatchPlan pp = patch_plan(sb, s.bwu);
            uint32_t        dn = 0, dextra = 0;
            const PatchPlan pd = delta_patch(x + kSampleVec[wave] * 1024, sb, s.bwdu, dn, dextra);
            if (lane == 0) {
                L.pbw[wave] = pp.bw, L.pexc[wave] = pp.exc;
                L.dpbw[wave] = pd.bw, L.dpexc[wave] = dn;
                L.bwf[wave] = s.bwf, L.bwd[wave] = s.bwd, L.runs[wave] = s.runs, L.bwp[wave] = s.bwp;
                L.eq[wave] = s.eq, L.nondec[wave] = s.nondec, L.inc[wave] = s.inc;
                L.umin[wave] = s.umin, L.umax[wave] = s.umax, L.smin[wave] = s.smin, L.smax[wave] = s.smax;
            }
        }
        __syncthreads();  // the sort buffers become the sample values in position order
        for (unsigned i = t; i < kFHt; i += blockDim.x) L.tab[i] = ~0u;
#pragma unroll
        for (unsigned j = 0; j < kSample / 256; ++j) {
            const unsigned p = t + 256 * j;
            L.u.xs[p]        = x[kSampleVec[p >> 10] * 1024 + (p & 1023)];
        }
        __syncthreads();
#pragma unroll
        for (unsigned j = 0; j < kSample / 256; ++j) {
            const unsigned p = t + 256 * j;
            L.slot[p]        = (uint16_t)pos_insert(L.tab, L.u.xs, p);
        }
        __syncthreads();
#pragma unroll
        for (unsigned j = 0; j < kSample / 256; ++j) {
            const unsigned p = t + 256 * j;
            if (L.tab[L.slot[p]] == p) atomicOr(&L.bitmap[p >> 5], 1u << (p & 31));
        }
        __syncthreads();
        const uint32_t D = block_popcount_scan(L.bitmap, L.base, kSample / 32, L.tmp);
        uint32_t       fp[kSample / 256];
#pragma unroll
        for (unsigned j = 0; j < kSample / 256; ++j) {
            const unsigned p  = t + 256 * j;
            fp[j]             = L.tab[L.slot[p]];
            const uint32_t c  = rank_of(L.bitmap, L.base, fp[j]);
            const uint32_t cp = (p % 32) ? rank_of(L.bitmap, L.base, L.tab[L.slot[p - 1]]) : c;
            const uint32_t d  = c - cp;
            const uint32_t cmn = wave_min(c), cmx = wave_max(c), dmn = wave_min(d), dmx = wave_max(d);
            if (lane == 0) {
                const unsigned v = p >> 10;
                atomicMin(&L.cmin[v], cmn), atomicMax(&L.cmax[v], cmx);
                atomicMin(&L.cdmin[v], dmn), atomicMax(&L.cdmax[v], dmx);
            }
        }
        __syncthreads();
        // multiplicities: the table's slots become counters
        for (unsigned i = t; i < kFHt; i += blockDim.x) L.tab[i] = 0;
        __syncthreads();
#pragma unroll
        for (unsigned j = 0; j < kSample / 256; ++j) atomicAdd(&L.tab[L.slot[t + 256 * j]], 1u);
        __syncthreads();
        {
            uint32_t m = 0;
            for (unsigned i = t; i < kFHt; i += blockDim.x) m = L.tab[i] > m ? L.tab[i] : m;
            m = wave_max(m);
            if (lane == 0) atomicMax(&L.top, m);
        }
        __syncthreads();
        if (t == 0) {
            float*         F  = feat + r * 64;
            const float    S  = (float)kSampleV, vv = (float)kVPR;
            uint32_t       sbwf = 0, sbwd = 0, sruns = 0, sbwp = 0, sbwc = 0, sbwcd = 0, eq = 0, nondec = 0, inc = 0;
            uint32_t       mxbwf = 0, mnbwf = 64, mxbwd = 0, mnbwd = 64, mxr = 0, mnr = 1024, spbw = 0, sexc = 0;
            uint32_t       sdpbw = 0, sdexc = 0;
            uint64_t       umax = 0, umin = kEmpty;
            int64_t        smax = INT64_MIN, smin = INT64_MAX;
            uint32_t       bwc[kSampleV], bwcd[kSampleV];
            for (unsigned i = 0; i < kSampleV; ++i) {
                bwc[i]  = bitlen64(L.cmax[i] - L.cmin[i]);
                bwcd[i] = bitlen64(L.cdmax[i] - L.cdmin[i]);
                sbwf += L.bwf[i], sbwd += L.bwd[i], sruns += L.runs[i], sbwp += L.bwp[i];
                sbwc += bwc[i], sbwcd += bwcd[i];
                spbw += L.pbw[i], sexc += L.pexc[i];
                sdpbw += L.dpbw[i], sdexc += L.dpexc[i];
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
            const float fD = (float)D, mpbw = (float)spbw / S, mexc = (float)sexc / S;
            const float mdpbw = (float)sdpbw / S, mdexc = (float)sdexc / S;
            float       est[kK];
            est[0] = 8.f + 4.f * vv + 8192.f * vv;
            est[1] = 8.f + 12.f * vv + vv * (9.f + 128.f * mbwf);
            est[2] = 8.f + 16.f * vv + vv * (137.f + 128.f * mbwd);
            est[3] = 8.f + 20.f * vv + vv * (8.f * mruns + 131.f + 128.f * mbwp);
            est[4] = 8.f + 8.f * vv + vv * (8.f * mruns + 2048.f);
            est[5] = 28.f + 8.f * fD + 4.f * vv + 4096.f * vv;
            est[6] = 28.f + 8.f * fD + 12.f * vv + vv * (5.f + 128.f * mbwc);
            est[7] = 28.f + 8.f * fD + 16.f * vv + vv * (133.f + 128.f * mbwcd);
            const float pk = 9.f + 128.f * mpbw;  // patched FFOR: bw, base and packed bytes per vector
            est[8]  = 8.f + 20.f * vv + vv * (10.f * mexc + 2.f + pk);
            est[9]  = 8.f + 20.f * vv + vv * (2.f * mexc + 2.f + pk) + 12.f + 8.f * mexc * vv;
            est[10] = 8.f + 20.f * vv + vv * (8.f * mexc + 128.f + pk);
            est[11] = 8.f + 20.f * vv + vv * (8.f * mexc + 4.f + pk);
            est[12] = 28.f + 8.f * fD + 20.f * vv + vv * (131.f + 128.f * mbwp) + 12.f + 4.f * mruns * vv;
            const float pkd = 9.f + 128.f * mdpbw;  // DELTA patches: bw, base and packed bytes per vector
            est[13] = 8.f + 24.f * vv + vv * (10.f * mdexc + 130.f + pkd);
            est[14] = 8.f + 24.f * vv + vv * (8.f * mdexc + 256.f + pkd);
            est[15] = 8.f + 24.f * vv + vv * (8.f * mdexc + 132.f + pkd);
            for (unsigned k = 0; k < kK; ++k) F[k] = est[k] / (float)(kRG * 8), F[kK + k] = log2f_(est[k]) / 24.f;
            float* G = F + 2 * kK;  // 32
            G[0] = mbwf / 64.f, G[1] = (float)mxbwf / 64.f, G[2] = (float)mnbwf / 64.f;
            G[3] = mbwd / 64.f, G[4] = (float)mxbwd / 64.f, G[5] = (float)mnbwd / 64.f;
            G[6] = mruns / 1024.f, G[7] = (float)mxr / 1024.f, G[8] = (float)mnr / 1024.f;
            G[9]  = log2f_(mruns) / 10.f;
            G[10] = mbwp;
            G[11] = fD / (S * 1024.f);
            G[12] = log2f_(fD) / 12.f;
            G[13] = (float)bitlen64(D - 1) / 16.f;
            G[14] = (float)L.top / (S * 1024.f);
            G[15] = (float)eq / (S * 1023.f);
            G[16] = (float)nondec / (S * 1023.f);
            G[17] = (float)inc / (S * 1023.f);
            G[18] = (float)bitlen64(umax) / 64.f;
            G[19] = (float)bitlen64(umin) / 64.f;
            G[20] = (float)bitlen64((uint64_t)smax - (uint64_t)smin) / 64.f;
            G[21] = mbwc / 32.f;
            G[22] = mbwcd / 32.f;
            for (unsigned i = 0; i < kSampleV; ++i) G[23 + i] = (float)L.bwf[i] / 64.f;  // 55 .. 58
            F[59] = mpbw / 64.f;
            F[60] = mexc / 100.f;
            F[61] = mdpbw / 64.f;
            F[62] = mdexc / 100.f;
            F[63] = 1.f;
        }
        __syncthreads();
    }
}

// per-device workspace of the exhaustive sizer
struct SizerWs {
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
    if (!w.firsts) {
        hipDeviceProp_t prop;
        if ((e = hipGetDeviceProperties(&prop, dev)) != hipSuccess) return hip_fail(e, "device properties");
        w.blocks = (unsigned)prop.multiProcessorCount;
        const size_t fb = (size_t)w.blocks * (2 * kHt + 64) * 4, sb = (size_t)w.blocks * kRG * 4;
        if ((e = hipMalloc(&w.firsts, fb)) != hipSuccess) return hip_fail(e, "hipMalloc");
        if ((e = hipMalloc(&w.slots, sb)) != hipSuccess) return hip_fail(e, "hipMalloc");
        // empty slots; the kernel restores this state after every row-group
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
    hipLaunchKernelGGL(k_wizard_sizes, dim3(blocks), dim3(1024), sizeof(SizerLds), (hipStream_t)stream, (const uint64_t*)dev_values,
                       n_rowgroups, w->firsts, w->slots, dev_sizes);
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
    hipLaunchKernelGGL(k_wizard_features, dim3(blocks), dim3(256), sizeof(FeatLds), (hipStream_t)stream,
                       (const uint64_t*)dev_values, n_rowgroups, dev_features);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "wizard features launch");
}

}  // extern "C"
