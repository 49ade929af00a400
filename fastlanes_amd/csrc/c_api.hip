// c_api.hip -- include/fastlanes.h implemented over the HIP decoder.
//
// Mirrors reference src/c_api/c_api_connector.cpp:6-37 and src/connection.cpp:152-180 (Connection's static
// encode_from_memory / decode_to_memory), with the behaviour changes listed in include/fastlanes.h.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/fastlanes.h"
#include "../../include/fls_gpu.h"
#include "format.hpp"
#include "gpu_internal.hpp"

using namespace fls;

struct Connection {
    hipStream_t stream = nullptr;
};

namespace {

bool is_device_ptr(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

// Bytes readable from p when p lies inside a device allocation (0 if unknown).
uint64_t device_extent(const void* p) {
    hipDeviceptr_t base = nullptr;
    size_t         size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return (uint64_t)((const uint8_t*)base + size - (const uint8_t*)p);
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// Decodes an encoded column whose bytes are at `in` (host or device) into `out` (host or device).
int decode_any(const uint8_t* in, uint64_t in_limit, void* out, uint8_t type, uint64_t* n_values_out) {
    const bool           in_dev = is_device_ptr(in);
    std::vector<uint8_t> host_copy;
    const uint8_t*       host = in;
    if (in_dev) {
        const uint64_t ext = device_extent(in);
        if (!ext) return fail(FLS_E_ARG, "cannot size the device input");
        host_copy.resize(ext);
        hipError_t e = hipMemcpy(host_copy.data(), in, ext, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
        host     = host_copy.data();
        in_limit = ext;
    }
    fls_gpu_plan* plan = nullptr;
    int           rc   = fls_gpu_plan_create(host, in_limit, (enum data_t)type, &plan);
    if (rc) return rc;
    const uint64_t n     = fls_gpu_plan_n_values(plan);
    const uint64_t bytes = n * (dtype_bits(type) / 8);
    const uint64_t size  = fls_gpu_plan_encoded_size(plan);
    DevBuf         din, dout;
    const uint8_t* d_in = in;
    hipError_t     e    = hipSuccess;
    if (!in_dev) {
        if ((e = hipMalloc(&din.p, size ? size : 1)) != hipSuccess) return fls_gpu_plan_destroy(plan), hip_fail(e, "hipMalloc");
        if ((e = hipMemcpy(din.p, in, size, hipMemcpyHostToDevice)) != hipSuccess)
            return fls_gpu_plan_destroy(plan), hip_fail(e, "hipMemcpy");
        d_in = (const uint8_t*)din.p;
    }
    const bool out_direct = is_device_ptr(out) && ((uintptr_t)out & 15) == 0;
    void*      d_out      = out;
    if (!out_direct) {
        if ((e = hipMalloc(&dout.p, bytes ? bytes : 16)) != hipSuccess) return fls_gpu_plan_destroy(plan), hip_fail(e, "hipMalloc");
        d_out = dout.p;
    }
    rc = fls_gpu_plan_decode(plan, d_in, d_out, bytes, nullptr);
    if (!rc) rc = fls_gpu_plan_check(plan, nullptr);
    fls_gpu_plan_destroy(plan);
    if (rc) return rc;
    if (!out_direct && bytes) {
        if ((e = hipMemcpy(out, d_out, bytes, hipMemcpyDefault)) != hipSuccess) return hip_fail(e, "hipMemcpy");
    }
    if (n_values_out) *n_values_out = n;
    return FLS_OK;
}

}  // namespace

extern "C" {

connection_cpi* fls_cpi_connection_open(void) { return new Connection(); }

void fls_cpi_connection_close(connection_cpi* c) { delete c; }

bool encode_from_memory(void* in_data, uint64_t in_n_input, uint64_t in_capacity, uint8_t* out_encoded_data,
                        uint64_t* out_encoded_size, uint64_t* out_metadadata_size, enum data_t in_datatype,
                        uint8_t in_expression_type) {
    if (!in_data || !out_encoded_data || !out_encoded_size) return fail(FLS_E_ARG, "null argument"), false;
    uint64_t meta = 0;
    int64_t  r    = encode_column((uint8_t)in_datatype, in_expression_type, in_data, in_n_input, out_encoded_data,
                                  in_capacity, &meta);
    if (r == -2) return fail(FLS_E_CAPACITY, "encoded column exceeds capacity"), false;
    if (r < 0) return fail(FLS_E_UNSUPPORTED, "unsupported type/expression or length not a multiple of 1024"), false;
    *out_encoded_size = (uint64_t)r;
    if (out_metadadata_size) *out_metadadata_size = meta;
    return true;
}

void decode_to_memory(uint8_t* in_encoded_data, void* out_data, enum data_t in_datatype) {
    if (!in_encoded_data || !out_data) {
        fail(FLS_E_ARG, "null argument");
        return;
    }
    // The reference signature carries no sizes: the column extent comes from its own page headers.
    int rc = decode_any(in_encoded_data, ~0ull - (uintptr_t)in_encoded_data, out_data, (uint8_t)in_datatype, nullptr);
    if (rc) std::fprintf(stderr, "fastlanes decode_to_memory: %s\n", fls_gpu_error_string());
}

// The reference's C ABI carries 64-bit columns only (INT64 / DOUBLE, src/encoder/single_col_encoder.cpp:13-33) and
// random access has no type argument, so the two functions below read 64-bit columns.
static uint64_t data_page_pos(const uint8_t* in) {
    const uint8_t first = in[0];
    if (first == EXP_FSST) {
        uint64_t tl = 17;
        for (int i = 0; i < 8; ++i) tl += (uint64_t)in[8 + 9 + i] * (i + 1);
        return 8 + tl;
    }
    if (!is_dict_exp(first)) return 0;
    uint64_t i = 8;  // HZL chain of 64-bit values (src/cor/eng/hzl_de_cmp.cpp:28-77)
    for (;;) {
        const uint8_t bw = in[i];
        uint16_t      exc;
        uint64_t      base;
        std::memcpy(&exc, in + i + 2, 2);
        std::memcpy(&base, in + i + 4, 8);
        i += 12;
        if (bw == 255) return i + base;
        i += 128ull * bw + (uint64_t)exc * (2 + 8);
    }
}

void init_random_access(uint8_t* in, uint64_t vec_id, uint64_t* n_buf, uint64_t* pairs) {
    if (!in || !n_buf || !pairs) return;
    *n_buf             = 0;
    const uint64_t pos = data_page_pos(in);
    const uint8_t* pg  = in + pos;
    Expr           e;
    if (!expr_from_page(pg[0], pg[1], e)) return;
    // buffers stored as one HZL chain per page (DICT_RLE, S_PATCH_REC, the RLE and B_PATCH_REC chain forms): no
    // vector owns a slice of them
    if (expr_rec(e)) return;
    const unsigned vec_c = (unsigned)pg[4] | ((unsigned)pg[5] << 8);
    if (vec_id >= vec_c) return;
    const unsigned v = (unsigned)vec_id;
    for (unsigned b = 0; b < e.arr_c; ++b) {
        const uint64_t s = off_of(pg, b) + ep_of(pg, e.arr_c, v, b);
        uint64_t       end;
        if (v + 1 < vec_c) end = off_of(pg, b) + ep_of(pg, e.arr_c, v + 1, b);
        else if (b + 1 < e.arr_c) end = off_of(pg, b + 1);
        else {  // last vector's last buffer: size from its bit width (one byte at a known offset)
            uint64_t tail = 0;
            switch (e.exp_t) {
            case EXP_UNCOMPRESSED: tail = 8192; break;
            case EXP_FSST: tail = 4096; break;
            case EXP_RLE: tail = e.var == 1 || e.var == kRleRec ? 2048 : 128ull * pg[off_of(pg, 2) + ep_of(pg, e.arr_c, v, 2)]; break;
            default: tail = e.var == 1 && is_dict_exp(e.exp_t) ? 4096 : 128ull * pg[off_of(pg, b - 2) + ep_of(pg, e.arr_c, v, b - 2)];
            }
            end = s + tail;
        }
        pairs[2 * b]     = pos + s;
        pairs[2 * b + 1] = end - s;
    }
    *n_buf = e.arr_c;
}

void random_access(uint8_t* in, uint64_t vec_id, uint64_t* n_buf, uint64_t* pairs, uint8_t** bufs, uint8_t* out_vec) {
    if (!in || !n_buf || !pairs || !bufs || !out_vec) return;
    const uint64_t pos = data_page_pos(in);
    const uint8_t* pg  = in + pos;
    Expr           e;
    if (!expr_from_page(pg[0], pg[1], e) || *n_buf != e.arr_c) {
        fail(FLS_E_FORMAT, "random_access: buffer count mismatch");
        return;
    }
    (void)vec_id;
    // A one-vector page (after the dictionary page, if any) rebuilt from the caller's buffers, decoded on the GPU.
    std::vector<uint8_t> mini(in, in + pos);
    const uint8_t        hdr[8] = {pg[0], pg[1], 0, 0, 1, 0, 0, 0};
    mini.insert(mini.end(), hdr, hdr + 8);
    uint32_t off = 8 + 4 * e.arr_c;
    for (unsigned b = 0; b < e.arr_c; ++b) {
        mini.insert(mini.end(), (uint8_t*)&off, (uint8_t*)&off + 4);
        off += (uint32_t)pairs[2 * b + 1];
    }
    for (unsigned b = 0; b < e.arr_c; ++b) mini.insert(mini.end(), bufs[b], bufs[b] + pairs[2 * b + 1]);
    int rc = decode_any(mini.data(), mini.size(), out_vec, DT_U64, nullptr);
    if (rc) std::fprintf(stderr, "fastlanes random_access: %s\n", fls_gpu_error_string());
}

}  // extern "C"
