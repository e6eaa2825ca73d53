// LDS ceilings of the device, for the rooflines of the LDS-bound kernels: the read rate
// of the interpolation kernels (512-thread workgroups read a 64 KB buffer with
// ds_read_b128, each lane 16 consecutive bytes, so a 64-lane read touches every bank
// evenly) and the 64-bit atomic-add rate of the spreading kernel.  Every index is
// masked into its buffer.
#include "common.h"

namespace {

constexpr int kProbeThreads = 512;
constexpr unsigned kProbeQuads = 4096;  // float4 entries: 64 KB

__global__ void __launch_bounds__(kProbeThreads) lds_read_probe_kernel(float* __restrict__ out, int iters) {
    __shared__ float4 buf[kProbeQuads];
    for (unsigned i = threadIdx.x; i < kProbeQuads; i += kProbeThreads)
        buf[i] = make_float4((float)i, (float)i + 1.f, (float)i + 2.f, (float)i + 3.f);
    __syncthreads();
    float acc = 0.f;
    for (int it = 0; it < iters; ++it) {
        const float4 v = buf[(threadIdx.x + (unsigned)it * 64u) & (kProbeQuads - 1)];
        acc += (v.x + v.y) + (v.z + v.w);
    }
    out[(size_t)blockIdx.x * kProbeThreads + threadIdx.x] = acc;  // keeps the reads
}

// 64-bit integer LDS atomics (the spreading kernel's accumulator adds): every lane adds
// to its own 8-byte counter, lanes of a wave on consecutive counters; indices masked
// into the 64 KB buffer.
constexpr unsigned kProbeCounters = 8192;  // 64 KB of uint64

__global__ void __launch_bounds__(kProbeThreads) lds_atomic_probe_kernel(float* __restrict__ out, int iters) {
    __shared__ unsigned long long acc[kProbeCounters];
    for (unsigned i = threadIdx.x; i < kProbeCounters; i += kProbeThreads) acc[i] = 0ull;
    __syncthreads();
    for (int it = 0; it < iters; ++it)
        atomicAdd(acc + ((threadIdx.x + (unsigned)it * 64u) & (kProbeCounters - 1)), (unsigned long long)(it + 1));
    __syncthreads();
    out[(size_t)blockIdx.x * kProbeThreads + threadIdx.x] = (float)acc[threadIdx.x];  // keeps the adds
}

}  // namespace

extern "C" int cmtip_lds_atomic_probe(int workgroups, int iters, float* out, void* stream) {
    if (workgroups < 1 || iters < 1 || !out) {
        ::cmtip::set_error("lds_atomic_probe: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipLaunchKernelGGL(lds_atomic_probe_kernel, dim3((unsigned)workgroups), dim3(kProbeThreads), 0,
                       ::cmtip::as_stream(stream), out, iters);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_lds_read_probe(int workgroups, int iters, float* out, void* stream) {
    if (workgroups < 1 || iters < 1 || !out) {
        ::cmtip::set_error("lds_read_probe: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipLaunchKernelGGL(lds_read_probe_kernel, dim3((unsigned)workgroups), dim3(kProbeThreads), 0,
                       ::cmtip::as_stream(stream), out, iters);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}
