// Shared helpers for the cmtip_amd HIP library (gfx950 / MI355X).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/cmtip_amd.h"

namespace cmtip {

// Per-thread message describing the last failure; read through cmtip_last_error().
void set_error(const char* fmt, ...);

#define CMTIP_HIP(call)                                                                 \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::cmtip::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,             \
                               hipGetErrorString(e_));                                  \
            return e_ == hipErrorOutOfMemory ? CMTIP_ERR_ALLOC : CMTIP_ERR_HIP;          \
        }                                                                               \
    } while (0)

#define CMTIP_LAUNCH_CHECK()                                                            \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess) {                                                         \
            ::cmtip::set_error("%s:%d launch -> %s", __FILE__, __LINE__,                \
                               hipGetErrorString(e_));                                  \
            return CMTIP_ERR_LAUNCH;                                                    \
        }                                                                               \
    } while (0)

#define CMTIP_TRY(expr)                \
    do {                               \
        int s_ = (expr);               \
        if (s_ != CMTIP_OK) return s_; \
    } while (0)

// Exact unsigned division by a runtime divisor d for x * d < 2^32:
// q = umulhi(x, ceil(2^32 / d)).
struct FastDiv {
    uint32_t d, m;
};
inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    f.m = (uint32_t)((((uint64_t)1 << 32) + d - 1) / d);
    if (d == 1) f.m = 0;  // handled by the d == 1 branch below
    return f;
}
__device__ __forceinline__ uint32_t fdiv(uint32_t x, FastDiv f) {
    return f.d == 1 ? x : __umulhi(x, f.m);
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Device buffer owned by a plan; freed by release().
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t nbytes) {
        if (nbytes <= bytes && p) return CMTIP_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (nbytes == 0) return CMTIP_OK;
        CMTIP_HIP(hipMalloc(&p, nbytes));
        bytes = nbytes;
        return CMTIP_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return reinterpret_cast<T*>(p); }
};

}  // namespace cmtip
