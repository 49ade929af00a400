// gpu_internal.hpp -- shared host-side pieces of the HIP layer.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "format.hpp"

namespace fls {

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Device-side FSST symbol table (include/fls/cor/prm/fsst/fsst.h:92-98 without the version/terminator fields).
struct FsstTableDev {
    uint64_t symbol[256];
    uint8_t  len[256];
};

struct StringPlan {
    FsstTableDev* d_tables = nullptr;  // one per page entry
    uint64_t*     d_status = nullptr;  // look-back status word per half-vector segment (zeroed once)
    uint64_t*     d_ticket = nullptr;  // look-back tickets: counter (never reset), the call's base, its epoch
    uint32_t*     d_err    = nullptr;  // error bits: capacity, look-back
    uint64_t      n_vec    = 0;
};

int  strings_prepare(const Column& col, const uint8_t* host, StringPlan& sp);
void strings_release(StringPlan& sp);
int  strings_offsets(const Column& col, StringPlan& sp, const uint8_t* dev_enc, uint64_t* dev_offsets, hipStream_t s);
int  strings_check(StringPlan& sp, hipStream_t s);
int  strings_decode(const Column& col, StringPlan& sp, const uint8_t* dev_enc, const uint64_t* dev_offsets,
                    uint8_t* dev_bytes, uint64_t cap, hipStream_t s);
int  strings_decode_all(const Column& col, StringPlan& sp, const uint8_t* dev_enc, uint64_t* dev_offsets,
                        uint8_t* dev_bytes, uint64_t cap, hipStream_t s);

}  // namespace fls
