// Shared glue of the C-ABI translation units (capi.hip, hooks.hip): the opaque handle
// types of include/stwo_rocm.h and the wrapper that locks a context and turns C++
// exceptions into status codes.
#pragma once
#include <string.h>

#include <functional>

#include "../../include/stwo_rocm.h"
#include "prover.hpp"

using namespace sr;

struct stwo_rocm_ctx {
    Ctx c;
    ServePool* pool = nullptr;  // serving slots, created by the first stwo_rocm_serve
};
struct stwo_rocm_tree {
    DeviceTree* t;
};
struct stwo_rocm_batch_prover {
    BatchPlan* p = nullptr;
    uint32_t n = 0;
};

template <class F>
static inline int guarded(stwo_rocm_ctx* ctx, F&& f) {
    if (!ctx) return STWO_ROCM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    try {
        SR_HIP(hipSetDevice(ctx->c.device));
        f(ctx->c);
        return STWO_ROCM_OK;
    } catch (const Error& e) {
        ctx->c.last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        ctx->c.last_error = e.what();
        return STWO_ROCM_EINTERNAL;
    }
}

// Column and QM31 arrays are read with 16-byte loads.
static inline void check_aligned(const void* const* ptrs, uint32_t n, const char* what) {
    for (uint32_t i = 0; i < n; i++)
        SR_CHECK(((uintptr_t)ptrs[i] & 15) == 0, std::string(what) + ": device pointers must be 16-byte aligned");
}

static inline qm31 qin(const uint32_t* w) {
    for (int i = 0; i < 4; i++) SR_CHECK(w[i] < P, "non-canonical M31 word");
    return qload(w);
}

static inline double timed(Ctx& c, const std::function<void()>& f) {
    if (!c.ev0) {
        SR_HIP(hipEventCreate(&c.ev0));
        SR_HIP(hipEventCreate(&c.ev1));
    }
    SR_HIP(hipEventRecord(c.ev0, c.stream));
    f();
    SR_HIP(hipEventRecord(c.ev1, c.stream));
    c.sync();
    float ms = 0;
    SR_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    return ms;
}

