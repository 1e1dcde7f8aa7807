// The context behind the opaque rgpu_ctx handle of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

struct rgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;  // default-priority stream owned by the context
    int cu_count = 0;
    uint32_t trace_waves = 0;   // persistent ray cast: resident waves per SIMD (0 = default)
    uint32_t trace_refill = 0;  // persistent ray cast: idle lanes that trigger a refill (0 = default)
    uint32_t trace_treelet = 0;     // persistent stackless ray cast: records staged in LDS (0 = default)
    uint32_t trace_leaf_batch = 0;  // persistent ray cast: waiting lanes that trigger leaf tests (0 = default)
    std::string arch;
    std::mutex err_mu;
    std::string last_error;
};

namespace rgpu {
int fail(rgpu_ctx* ctx, int code, const std::string& msg);
int hip_fail(rgpu_ctx* ctx, hipError_t e, const char* what);
hipStream_t stream_of(rgpu_ctx* ctx, void* s);
// RGPU_VARIANT_STACKLESS (0), _THREADED (1) or _THREADED_SINGLE (2), which equal the kernels'
// kVariant* ids (checked in capi.cpp)
inline bool valid_variant(int v) { return v >= 0 && v <= 2; }
}  // namespace rgpu
