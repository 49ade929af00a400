// mlp.hip -- the learned encoding selector: a 64-256-256-K MLP trained and run on the GPU.
//
// GEMMs run on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation). Every GEMM is
// written C[M][N] = sum_k A[M][k] * B[N][k] with both operands K-contiguous, so the forward pass, the activation
// gradients and the weight gradients all use one kernel; the epilogue adds a bias, applies ReLU or a ReLU mask and
// stores FP32, BF16 and/or transposed BF16 results. Master weights are FP32 (Adam); BF16 copies in both layouts are
// published after every step. No counterpart exists in the reference (its wizard has no learned part,
// src/wizard/wizard.cpp:46-66 leaves the size-driven choice empty).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "../../include/fls_gpu.h"
#include "../../include/fls_wizard.h"
#include "gpu_internal.hpp"

namespace fls {
namespace {

constexpr unsigned kF = 64, kH = 256, kKP = 32;  // features, hidden, padded classes
constexpr uint64_t kOffW1 = 0, kOffB1 = kOffW1 + kH * kF, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
                   kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kKP * kH, kNParams = kOffB3 + kKP;

typedef __bf16   bf16x8 __attribute__((ext_vector_type(8)));
typedef float    f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct GemmArgs {
    const uint16_t* A;  // [M][lda] bf16
    const uint16_t* B;  // [N][ldb] bf16
    int             M, N, K, lda, ldb;
    const float*    bias;  // [N] or null
    int             relu;
    const uint16_t* mask;  // [M][ldm] bf16: output *= (mask > 0), or null
    int             ldm;
    float*          Cf;  // [M][ldcf] fp32 or null
    int             ldcf;
    uint16_t*       Cb;  // [M][ldcb] bf16 or null
    int             ldcb;
    uint16_t*       CbT;  // [N][ldct] bf16 (transposed) or null
    int             ldct;
};

// 256 threads = 4 waves; block tile 64x64, wave tile 32x32 = 2x2 MFMA 16x16x32 blocks. M, N multiples of 32,
// K a multiple of 32 (checked on the host); waves whose tile lies outside M x N return.
__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
    const unsigned lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int      m0 = blockIdx.y * 64 + (wave / 2) * 32, n0 = blockIdx.x * 64 + (wave % 2) * 32;
    if (m0 >= g.M || n0 >= g.N) return;
    const int r = lane % 16, kq = (lane / 16) * 8;
    f32x4     acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4 {0.f, 0.f, 0.f, 0.f};
    const uint16_t* a0 = g.A + (size_t)(m0 + r) * g.lda + kq;
    const uint16_t* a1 = a0 + (size_t)16 * g.lda;
    const uint16_t* b0 = g.B + (size_t)(n0 + r) * g.ldb + kq;
    const uint16_t* b1 = b0 + (size_t)16 * g.ldb;
    for (int k = 0; k < g.K; k += 32) {
        const bf16x8 fa0 = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(a0 + k));
        const bf16x8 fa1 = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(a1 + k));
        const bf16x8 fb0 = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(b0 + k));
        const bf16x8 fb1 = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(b1 + k));
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa0, fb0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa0, fb1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa1, fb0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa1, fb1, acc[1][1], 0, 0, 0);
    }
    // accumulator lane layout: element q of lane l is C[4 * (l / 16) + q][l % 16] of the 16x16 block
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + 16 * i + 4 * (lane / 16) + q, n = n0 + 16 * j + (lane % 16);
                float     v = acc[i][j][q];
                if (g.bias) v += g.bias[n];
                if (g.relu) v = v > 0.f ? v : 0.f;
                if (g.mask && !(from_bf16(g.mask[(size_t)m * g.ldm + n]) > 0.f)) v = 0.f;
                if (g.Cf) g.Cf[(size_t)m * g.ldcf + n] = v;
                const uint16_t h = to_bf16(v);
                if (g.Cb) g.Cb[(size_t)m * g.ldcb + n] = h;
                if (g.CbT) g.CbT[(size_t)n * g.ldct + m] = h;
            }
}

// features fp32 [B][64] -> bf16 [B][64] and transposed [64][B]
__global__ void k_cast_x(const float* __restrict__ X, uint32_t B, uint16_t* __restrict__ Xb, uint16_t* __restrict__ XbT) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)B * kF) return;
    const uint32_t b = (uint32_t)(i / kF), f = (uint32_t)(i % kF);
    const uint16_t h = to_bf16(X[i]);
    Xb[i]            = h;
    if (XbT) XbT[(uint64_t)f * B + b] = h;
}

// softmax cross-entropy over the first kc logits of each row; dZ = (softmax - onehot) / B
__global__ void k_softmax_xent(const float* __restrict__ Z, const int32_t* __restrict__ y, uint32_t B, uint32_t kc,
                               float* __restrict__ dZf, uint16_t* __restrict__ dZb, uint16_t* __restrict__ dZbT,
                               float* __restrict__ loss) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float* z  = Z + (uint64_t)b * kKP;
    float        mx = z[0];
    for (uint32_t k = 1; k < kc; ++k) mx = z[k] > mx ? z[k] : mx;
    float s = 0.f;
    for (uint32_t k = 0; k < kc; ++k) s += expf(z[k] - mx);
    int32_t lab = y[b];
    lab = lab < 0 ? 0 : (lab >= (int32_t)kc ? (int32_t)kc - 1 : lab);  // labels outside [0, kc) are clamped
    atomicAdd(loss, (logf(s) + mx - z[lab]) / (float)B);
    for (uint32_t k = 0; k < kKP; ++k) {
        float d = 0.f;
        if (k < kc) d = (expf(z[k] - mx) / s - (k == (uint32_t)lab ? 1.f : 0.f)) / (float)B;
        dZf[(uint64_t)b * kKP + k]  = d;
        const uint16_t h            = to_bf16(d);
        dZb[(uint64_t)b * kKP + k]  = h;
        dZbT[(uint64_t)k * B + b]   = h;
    }
}

// bias gradients: g[n] = sum_b dZf[b][n] (fp32 rows of ld), one wave per column ...
__global__ void k_colsum_f32(const float* __restrict__ src, uint32_t B, uint32_t ld, float* __restrict__ g) {
    const uint32_t n = blockIdx.x, lane = threadIdx.x;
    float          s = 0.f;
    for (uint32_t b = lane; b < B; b += 64) s += src[(uint64_t)b * ld + n];
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (lane == 0) g[n] = s;
}
// ... or g[n] = sum_b srcT[n][b] (bf16 transposed rows), one wave per row
__global__ void k_rowsum_bf16(const uint16_t* __restrict__ srcT, uint32_t B, float* __restrict__ g) {
    const uint32_t n = blockIdx.x, lane = threadIdx.x;
    float          s = 0.f;
    for (uint32_t b = lane; b < B; b += 64) s += from_bf16(srcT[(uint64_t)n * B + b]);
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (lane == 0) g[n] = s;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       uint64_t n, float lr, float c1, float c2) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi, vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi, v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// bf16 copies of the weights: W1t, W2t, W3t as stored ([out][in]) and W2, W3 transposed ([in][out])
__global__ void k_publish(const float* __restrict__ p, uint16_t* __restrict__ w1t, uint16_t* __restrict__ w2t,
                          uint16_t* __restrict__ w2, uint16_t* __restrict__ w3t, uint16_t* __restrict__ w3) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kH * kF) w1t[i] = to_bf16(p[kOffW1 + i]);
    if (i < kH * kH) {
        const uint16_t h = to_bf16(p[kOffW2 + i]);
        w2t[i]           = h;
        w2[(i % kH) * kH + i / kH] = h;
    }
    if (i < kKP * kH) {
        const uint16_t h = to_bf16(p[kOffW3 + i]);
        w3t[i]           = h;
        w3[(i % kH) * kKP + i / kH] = h;
    }
}

__global__ void k_argmax(const float* __restrict__ Z, uint32_t B, uint32_t kc, int32_t* __restrict__ out) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float* z    = Z + (uint64_t)b * kKP;
    int32_t      best = 0;
    for (uint32_t k = 1; k < kc; ++k)
        if (z[k] > z[best]) best = (int32_t)k;
    out[b] = best;
}

}  // namespace
}  // namespace fls

using namespace fls;

struct fls_mlp {
    uint32_t kc = 0, max_batch = 0, batch = 0, step = 0;
    float*   P = nullptr;  // params, grads, adam m, adam v (4 x kNParams)
    float *  G = nullptr, *Am = nullptr, *Av = nullptr;
    uint16_t *w1t = nullptr, *w2t = nullptr, *w2 = nullptr, *w3t = nullptr, *w3 = nullptr;
    uint16_t *Xb = nullptr, *XbT = nullptr, *H1 = nullptr, *H1T = nullptr, *H2 = nullptr, *H2T = nullptr;
    uint16_t *dZb = nullptr, *dZbT = nullptr, *dH2 = nullptr, *dH2T = nullptr, *dH1T = nullptr;
    float *   Z = nullptr, *dZf = nullptr, *loss = nullptr;
    std::vector<void*> allocs;
};

namespace {

template <typename T>
int alloc(fls_mlp* m, T*& p, size_t count) {
    void*      q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count * sizeof(T), 256));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    m->allocs.push_back(q);
    p = (T*)q;
    return FLS_OK;
}

int launch_gemm(const GemmArgs& g, hipStream_t s) {
    if (g.M % 32 || g.N % 32 || g.K % 32 || g.K == 0) return fail(FLS_E_ARG, "gemm shape must be multiples of 32");
    hipLaunchKernelGGL(k_gemm_bf16, dim3((g.N + 63) / 64, (g.M + 63) / 64), dim3(256), 0, s, g);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "gemm launch");
}

GemmArgs gemm(const uint16_t* A, int lda, const uint16_t* B, int ldb, int M, int N, int K) {
    GemmArgs g {};
    g.A = A, g.lda = lda, g.B = B, g.ldb = ldb, g.M = M, g.N = N, g.K = K;
    return g;
}

int publish(fls_mlp* m, hipStream_t s) {
    hipLaunchKernelGGL(k_publish, dim3((kH * kH + 255) / 256), dim3(256), 0, s, m->P, m->w1t, m->w2t, m->w2, m->w3t,
                       m->w3);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "publish launch");
}

int forward(fls_mlp* m, const float* X, uint32_t B, bool keep_t, hipStream_t s) {
    hipLaunchKernelGGL(k_cast_x, dim3((B * kF + 255) / 256), dim3(256), 0, s, X, B, m->Xb, keep_t ? m->XbT : nullptr);
    GemmArgs g = gemm(m->Xb, kF, m->w1t, kF, (int)B, kH, kF);
    g.bias = m->P + kOffB1, g.relu = 1, g.Cb = m->H1, g.ldcb = kH;
    if (keep_t) g.CbT = m->H1T, g.ldct = (int)B;
    int rc = launch_gemm(g, s);
    if (rc) return rc;
    g      = gemm(m->H1, kH, m->w2t, kH, (int)B, kH, kH);
    g.bias = m->P + kOffB2, g.relu = 1, g.Cb = m->H2, g.ldcb = kH;
    if (keep_t) g.CbT = m->H2T, g.ldct = (int)B;
    if ((rc = launch_gemm(g, s))) return rc;
    g      = gemm(m->H2, kH, m->w3t, kH, (int)B, kKP, kH);
    g.bias = m->P + kOffB3, g.Cf = m->Z, g.ldcf = kKP;
    return launch_gemm(g, s);
}

}  // namespace

extern "C" {

uint64_t fls_mlp_n_params(void) { return kNParams; }

int fls_mlp_create(uint32_t n_classes, uint32_t max_batch, uint64_t seed, fls_mlp** out) {
    if (!out) return fail(FLS_E_ARG, "null argument");
    *out = nullptr;
    if (n_classes < 2 || n_classes > kKP) return fail(FLS_E_ARG, "n_classes must be in [2, 32]");
    if (max_batch == 0 || max_batch % 32) return fail(FLS_E_ARG, "max_batch must be a positive multiple of 32");
    auto* m      = new fls_mlp();
    m->kc        = n_classes;
    m->max_batch = max_batch;
    const size_t Bm = max_batch;
    int          rc = 0;
    (void)(rc || (rc = alloc(m, m->P, kNParams)) || (rc = alloc(m, m->G, kNParams)) || (rc = alloc(m, m->Am, kNParams)) ||
           (rc = alloc(m, m->Av, kNParams)) || (rc = alloc(m, m->w1t, kH * kF)) || (rc = alloc(m, m->w2t, kH * kH)) ||
           (rc = alloc(m, m->w2, kH * kH)) || (rc = alloc(m, m->w3t, kKP * kH)) || (rc = alloc(m, m->w3, kKP * kH)) ||
           (rc = alloc(m, m->Xb, Bm * kF)) || (rc = alloc(m, m->XbT, Bm * kF)) || (rc = alloc(m, m->H1, Bm * kH)) ||
           (rc = alloc(m, m->H1T, Bm * kH)) || (rc = alloc(m, m->H2, Bm * kH)) || (rc = alloc(m, m->H2T, Bm * kH)) ||
           (rc = alloc(m, m->dZb, Bm * kKP)) || (rc = alloc(m, m->dZbT, Bm * kKP)) || (rc = alloc(m, m->dH2, Bm * kH)) ||
           (rc = alloc(m, m->dH2T, Bm * kH)) || (rc = alloc(m, m->dH1T, Bm * kH)) || (rc = alloc(m, m->Z, Bm * kKP)) ||
           (rc = alloc(m, m->dZf, Bm * kKP)) || (rc = alloc(m, m->loss, 1)));
    if (rc) {
        fls_mlp_destroy(m);
        return rc;
    }
    // He-normal weights, zero biases; logits of unused classes stay 0
    std::vector<float> p(kNParams, 0.f);
    std::mt19937_64    gen(seed);
    auto               fill = [&](uint64_t off, unsigned rows, unsigned cols, unsigned fan_in) {
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)fan_in));
        for (unsigned i = 0; i < rows * cols; ++i) p[off + i] = nd(gen);
    };
    fill(kOffW1, kH, kF, kF);
    fill(kOffW2, kH, kH, kH);
    fill(kOffW3, n_classes, kH, kH);
    rc = fls_mlp_set_params(m, p.data());
    hipError_t e = hipSuccess;
    if (!rc && (e = hipMemset(m->Am, 0, kNParams * 4)) != hipSuccess) rc = hip_fail(e, "hipMemset");
    if (!rc && (e = hipMemset(m->Av, 0, kNParams * 4)) != hipSuccess) rc = hip_fail(e, "hipMemset");
    if (!rc && (e = hipMemset(m->G, 0, kNParams * 4)) != hipSuccess) rc = hip_fail(e, "hipMemset");
    if (!rc && (e = hipDeviceSynchronize()) != hipSuccess) rc = hip_fail(e, "sync");
    if (rc) {
        fls_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return FLS_OK;
}

void fls_mlp_destroy(fls_mlp* m) {
    if (!m) return;
    for (void* q : m->allocs) (void)hipFree(q);
    delete m;
}

int fls_mlp_get_params(fls_mlp* m, float* host_params) {
    if (!m || !host_params) return fail(FLS_E_ARG, "null argument");
    hipError_t e = hipMemcpy(host_params, m->P, kNParams * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? FLS_OK : hip_fail(e, "hipMemcpy");
}

int fls_mlp_set_params(fls_mlp* m, const float* host_params) {
    if (!m || !host_params) return fail(FLS_E_ARG, "null argument");
    hipError_t e = hipMemcpy(m->P, host_params, kNParams * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
    int rc = publish(m, nullptr);
    if (rc) return rc;
    e = hipDeviceSynchronize();
    return e == hipSuccess ? FLS_OK : hip_fail(e, "sync");
}

int fls_mlp_get_grads(fls_mlp* m, float* host_grads) {
    if (!m || !host_grads) return fail(FLS_E_ARG, "null argument");
    hipError_t e = hipMemcpy(host_grads, m->G, kNParams * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? FLS_OK : hip_fail(e, "hipMemcpy");
}

int fls_mlp_forward_backward(fls_mlp* m, const float* dev_features, const int32_t* dev_labels, uint32_t batch,
                             float* host_loss, void* stream) {
    if (!m || !dev_features || !dev_labels) return fail(FLS_E_ARG, "null argument");
    if (batch == 0 || batch % 32 || batch > m->max_batch) return fail(FLS_E_ARG, "batch must be a multiple of 32 <= max_batch");
    hipStream_t s = (hipStream_t)stream;
    const int   B = (int)batch;
    m->batch      = batch;
    int rc = forward(m, dev_features, batch, true, s);
    if (rc) return rc;
    hipError_t e = hipMemsetAsync(m->loss, 0, 4, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
    hipLaunchKernelGGL(k_softmax_xent, dim3((batch + 255) / 256), dim3(256), 0, s, m->Z, dev_labels, batch, m->kc,
                       m->dZf, m->dZb, m->dZbT, m->loss);
    // output layer
    GemmArgs g = gemm(m->dZbT, B, m->H2T, B, kKP, kH, B);
    g.Cf = m->G + kOffW3, g.ldcf = kH;
    if ((rc = launch_gemm(g, s))) return rc;
    hipLaunchKernelGGL(k_colsum_f32, dim3(kKP), dim3(64), 0, s, m->dZf, batch, kKP, m->G + kOffB3);
    // hidden layer 2
    g      = gemm(m->dZb, kKP, m->w3, kKP, B, kH, kKP);
    g.mask = m->H2, g.ldm = kH, g.Cb = m->dH2, g.ldcb = kH, g.CbT = m->dH2T, g.ldct = B;
    if ((rc = launch_gemm(g, s))) return rc;
    g    = gemm(m->dH2T, B, m->H1T, B, kH, kH, B);
    g.Cf = m->G + kOffW2, g.ldcf = kH;
    if ((rc = launch_gemm(g, s))) return rc;
    hipLaunchKernelGGL(k_rowsum_bf16, dim3(kH), dim3(64), 0, s, m->dH2T, batch, m->G + kOffB2);
    // hidden layer 1
    g      = gemm(m->dH2, kH, m->w2, kH, B, kH, kH);
    g.mask = m->H1, g.ldm = kH, g.CbT = m->dH1T, g.ldct = B;
    if ((rc = launch_gemm(g, s))) return rc;
    g    = gemm(m->dH1T, B, m->XbT, B, kH, kF, B);
    g.Cf = m->G + kOffW1, g.ldcf = kF;
    if ((rc = launch_gemm(g, s))) return rc;
    hipLaunchKernelGGL(k_rowsum_bf16, dim3(kH), dim3(64), 0, s, m->dH1T, batch, m->G + kOffB1);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "backward launch");
    if (host_loss) {
        if ((e = hipMemcpyAsync(host_loss, m->loss, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return hip_fail(e, "copy");
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "sync");
    }
    return FLS_OK;
}

int fls_mlp_adam_step(fls_mlp* m, float lr, void* stream) {
    if (!m) return fail(FLS_E_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    m->step++;
    const float c1 = 1.f - std::pow(0.9f, (float)m->step), c2 = 1.f - std::pow(0.999f, (float)m->step);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((kNParams + 255) / 256)), dim3(256), 0, s, m->P, m->G, m->Am, m->Av,
                       kNParams, lr, c1, c2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "adam launch");
    return publish(m, s);
}

int fls_mlp_predict(fls_mlp* m, const float* dev_features, uint64_t n, int32_t* dev_choice, void* stream) {
    if (!m || (n && (!dev_features || !dev_choice))) return fail(FLS_E_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    for (uint64_t i = 0; i < n; i += m->max_batch) {
        const uint32_t rows = (uint32_t)std::min<uint64_t>(m->max_batch, n - i);
        const uint32_t B    = (rows + 31) / 32 * 32;  // rows past n read the zero-padded tail of Xb
        if (B != rows) {
            hipError_t e = hipMemsetAsync(m->Xb, 0, (size_t)B * kF * 2, s);
            if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
        }
        hipLaunchKernelGGL(k_cast_x, dim3((rows * kF + 255) / 256), dim3(256), 0, s, dev_features + i * kF, rows, m->Xb,
                           nullptr);
        GemmArgs g = gemm(m->Xb, kF, m->w1t, kF, (int)B, kH, kF);
        g.bias = m->P + kOffB1, g.relu = 1, g.Cb = m->H1, g.ldcb = kH;
        int rc = launch_gemm(g, s);
        if (rc) return rc;
        g      = gemm(m->H1, kH, m->w2t, kH, (int)B, kH, kH);
        g.bias = m->P + kOffB2, g.relu = 1, g.Cb = m->H2, g.ldcb = kH;
        if ((rc = launch_gemm(g, s))) return rc;
        g      = gemm(m->H2, kH, m->w3t, kH, (int)B, kKP, kH);
        g.bias = m->P + kOffB3, g.Cf = m->Z, g.ldcf = kKP;
        if ((rc = launch_gemm(g, s))) return rc;
        hipLaunchKernelGGL(k_argmax, dim3((rows + 255) / 256), dim3(256), 0, s, m->Z, rows, m->kc, dev_choice + i);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "predict launch");
    }
    return FLS_OK;
}

int fls_mlp_get_logits(fls_mlp* m, float* host_logits, uint32_t batch) {
    if (!m || !host_logits || batch > m->max_batch) return fail(FLS_E_ARG, "bad argument");
    hipError_t e = hipMemcpy(host_logits, m->Z, (size_t)batch * kKP * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? FLS_OK : hip_fail(e, "hipMemcpy");
}

}  // extern "C"
