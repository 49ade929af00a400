// rowgroup.hip -- row-group decode: many columns, one launch.
#include <memory>
#include <vector>

#include "../../include/fls_gpu.h"
#include "format.hpp"
#include "gpu_internal.hpp"

using namespace fls;

struct fls_gpu_rowgroup {
    std::vector<fls_gpu_plan*> plans;
    std::vector<uint64_t>      dev_off, out_off;
    std::vector<uint8_t>       types;
    uint64_t                   out_total = 0;
};

extern "C" {

int fls_gpu_rowgroup_create(uint32_t n_columns, const uint8_t* const* host_columns, const uint64_t* column_sizes,
                            const uint64_t* device_offsets, const enum data_t* types, fls_gpu_rowgroup** out) {
    if (!out || (n_columns && (!host_columns || !column_sizes || !device_offsets || !types)))
        return fail(FLS_E_ARG, "null argument");
    auto rg = std::make_unique<fls_gpu_rowgroup>();
    for (uint32_t c = 0; c < n_columns; ++c) {
        fls_gpu_plan* p  = nullptr;
        int           rc = fls_gpu_plan_create(host_columns[c], column_sizes[c], types[c], &p);
        if (rc) {
            for (auto* q : rg->plans) fls_gpu_plan_destroy(q);
            return rc;
        }
        if (types[c] == STR) {
            fls_gpu_plan_destroy(p);
            for (auto* q : rg->plans) fls_gpu_plan_destroy(q);
            return fail(FLS_E_UNSUPPORTED, "string columns are decoded with fls_gpu_plan_decode_strings");
        }
        rg->plans.push_back(p);
        rg->dev_off.push_back(device_offsets[c]);
        rg->types.push_back((uint8_t)types[c]);
        rg->out_off.push_back(rg->out_total);
        const uint64_t bytes = fls_gpu_plan_n_values(p) * (dtype_bits((uint8_t)types[c]) / 8);
        rg->out_total += (bytes + 255) & ~255ull;
    }
    *out = rg.release();
    return FLS_OK;
}

void fls_gpu_rowgroup_destroy(fls_gpu_rowgroup* rg) {
    if (!rg) return;
    for (auto* p : rg->plans) fls_gpu_plan_destroy(p);
    delete rg;
}

int fls_gpu_rowgroup_layout(const fls_gpu_rowgroup* rg, uint64_t* out_offsets, uint64_t* out_total_bytes) {
    if (!rg) return fail(FLS_E_ARG, "null argument");
    if (out_offsets)
        for (size_t c = 0; c < rg->out_off.size(); ++c) out_offsets[c] = rg->out_off[c];
    if (out_total_bytes) *out_total_bytes = rg->out_total;
    return FLS_OK;
}

int fls_gpu_rowgroup_decode(fls_gpu_rowgroup* rg, const uint8_t* dev_encoded, uint8_t* dev_out, int mode, void* stream) {
    if (!rg || !dev_encoded || !dev_out) return fail(FLS_E_ARG, "null argument");
    if (mode != 1) return fail(FLS_E_UNSUPPORTED, "persistent row-group launch not built");
    for (size_t c = 0; c < rg->plans.size(); ++c) {
        const uint64_t bytes = fls_gpu_plan_n_values(rg->plans[c]) * (dtype_bits(rg->types[c]) / 8);
        int rc = fls_gpu_plan_decode(rg->plans[c], dev_encoded + rg->dev_off[c], dev_out + rg->out_off[c], bytes, stream);
        if (rc) return rc;
    }
    return FLS_OK;
}

int fls_gpu_rowgroup_decode_graph(fls_gpu_rowgroup* rg, const uint8_t* dev_encoded, uint8_t* dev_out, void* stream) {
    (void)rg, (void)dev_encoded, (void)dev_out, (void)stream;
    return fail(FLS_E_UNSUPPORTED, "graph capture not built");
}

int fls_gpu_rowgroup_check(fls_gpu_rowgroup* rg, void* stream) {
    if (!rg) return fail(FLS_E_ARG, "null argument");
    for (auto* p : rg->plans) {
        int rc = fls_gpu_plan_check(p, stream);
        if (rc) return rc;
    }
    return FLS_OK;
}

}  // extern "C"
