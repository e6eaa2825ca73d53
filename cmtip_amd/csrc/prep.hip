// Data layer of the reconstruction: which pixels enter the hot path and at what binning
// (cmtip/prep_data.py:96-262).  HBM-bound integer / copy work: detector binning by block
// averaging, resolution clipping and trimming as order-preserving column gathers, the
// assembly of panels into a 2-D detector through the pixel index map, and the binning of
// that index map.
//
// Binning reproduces numpy's float32 arithmetic exactly (prep_data.py:216-221:
// reshape(..., X/b, b, Y/b, b).mean(-1).mean(3)): the inner mean sums b contiguous
// values (sequentially below 8 values, with numpy's eight interleaved partial sums from
// 8 to 128) and divides by b; the outer mean adds those b row means sequentially and
// divides by b, every operation rounded to float32.
#include <stdint.h>

#include "common.h"

using namespace cmtip;

namespace {

constexpr int kT = 256;

unsigned grid_for(int64_t n, int64_t cap = 65535 * 4) {
    int64_t b = (n + kT - 1) / kT;
    if (b > cap) b = cap;
    return (unsigned)(b < 1 ? 1 : b);
}

// sum of n contiguous float32 values in numpy's order (pairwise_sum for n <= 128)
__device__ __forceinline__ float np_sum_f32(const float* __restrict__ v, int n) {
#pragma clang fp contract(off)
    if (n < 8) {
        float s = v[0];
        for (int i = 1; i < n; ++i) s += v[i];
        return s;
    }
    float r[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = v[k];
    int i = 8;
    for (; i + 8 <= n; i += 8)
#pragma unroll
        for (int k = 0; k < 8; ++k) r[k] += v[i + k];
    float s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) s += v[i];
    return s;
}

// out[s, i, j] = mean over the b x b block (rows b i .. b i + b - 1, columns b j ..) of
// in[s], with numpy's order of operations (see the file comment)
__global__ void bin_images_kernel(const float* __restrict__ in, int64_t n_stack, int nx, int ny, int b,
                                  float* __restrict__ out) {
#pragma clang fp contract(off)
    const int ox = nx / b, oy = ny / b;
    const int64_t total = n_stack * ox * oy;
    const float fb = (float)b;
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
        const int j = (int)(t % oy);
        const int64_t r = t / oy;
        const int i = (int)(r % ox);
        const int64_t s = r / ox;
        const float* base = in + (s * nx + (int64_t)i * b) * ny + (int64_t)j * b;
        float acc = np_sum_f32(base, b) / fb;
        for (int k = 1; k < b; ++k) acc += np_sum_f32(base + (int64_t)k * ny, b) / fb;
        out[t] = acc / fb;
    }
}

// flags[p] = resolution of pixel p > res_limit with the reference's float32 arithmetic
// (prep_data.py:166-167): norm = sqrt((x^2 + y^2) + z^2), res = (1 / norm) * 1e10, all
// rounded to float32, compared with res_limit rounded to float32
__global__ void resolution_flags_kernel(const float* __restrict__ s, int64_t stride, int64_t P, float lim,
                                        uint8_t* __restrict__ flags) {
#pragma clang fp contract(off)
    for (int64_t p = (int64_t)blockIdx.x * kT + threadIdx.x; p < P; p += (int64_t)gridDim.x * kT) {
        const float x = s[p], y = s[stride + p], z = s[2 * stride + p];
        const float n2 = (x * x + y * y) + z * z;
        const float res = (1.0f / sqrtf(n2)) * 1e10f;
        flags[p] = res > lim ? 1 : 0;
    }
}

// order-preserving compaction of flags into indices; one workgroup of 1024 threads
// walking the flags in tiles with a running offset
constexpr int kScan = 1024;
__global__ void __launch_bounds__(kScan) compact_kernel(const uint8_t* __restrict__ flags, int64_t n,
                                                        int64_t* __restrict__ idx, int64_t* __restrict__ count) {
    __shared__ int64_t sc[kScan];
    int64_t base = 0;
    for (int64_t t0 = 0; t0 < n; t0 += kScan) {
        const int64_t t = t0 + threadIdx.x;
        const int f = t < n ? flags[t] : 0;
        sc[threadIdx.x] = f;
        __syncthreads();
        for (int off = 1; off < kScan; off <<= 1) {  // inclusive Hillis-Steele scan
            const int64_t v = threadIdx.x >= (unsigned)off ? sc[threadIdx.x - off] : 0;
            __syncthreads();
            sc[threadIdx.x] += v;
            __syncthreads();
        }
        if (f) idx[base + sc[threadIdx.x] - 1] = t;
        const int64_t tot = sc[kScan - 1];
        __syncthreads();
        base += tot;
    }
    if (threadIdx.x == 0) *count = base;
}

// out[r, k] = in[r, idx[k]] (0 where idx[k] < 0): rows of P_in values in, P_out out
template <class T>
__global__ void gather_cols_kernel(const T* __restrict__ in, int64_t rows, int64_t P_in,
                                   const int64_t* __restrict__ idx, int64_t P_out, T* __restrict__ out) {
    const int64_t total = rows * P_out;
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
        const int64_t r = t / P_out, k = t - r * P_out;
        const int64_t src = idx[k];
        out[t] = src >= 0 ? in[r * P_in + src] : T(0);
    }
}

__global__ void fill_i64_kernel(int64_t* __restrict__ p, int64_t n, int64_t v) {
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < n; t += (int64_t)gridDim.x * kT) p[t] = v;
}

// src[X * Y] = the largest flat pixel index (l * px * py + i * py + j) that the index map
// sends to (X, Y): numpy's fancy assignment of assemble_image_stack_batch
// (prep_data.py:251-252) leaves the last pixel written there, panels in order
__global__ void assemble_src_kernel(const int64_t* __restrict__ imap, int64_t n_pix, int64_t X, int64_t Y,
                                    unsigned long long* __restrict__ src) {
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < n_pix; t += (int64_t)gridDim.x * kT) {
        const int64_t ix = imap[2 * t], iy = imap[2 * t + 1];
        if (ix >= 0 && ix < X && iy >= 0 && iy < Y) atomicMax(src + ix * Y + iy, (unsigned long long)(t + 1));
    }
}

__global__ void src_to_index_kernel(const unsigned long long* __restrict__ src, int64_t n,
                                    int64_t* __restrict__ idx) {
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < n; t += (int64_t)gridDim.x * kT)
        idx[t] = (int64_t)src[t] - 1;
}

// bin_pixel_index_map (prep_data.py:229-242): out[l, i, j, c] = min of in[l, b i + a,
// b j + e, c] over a, e in {0, 1}, floor-divided by b
__global__ void bin_index_map_kernel(const int64_t* __restrict__ in, int64_t n_panels, int px, int py, int b,
                                     int64_t* __restrict__ out) {
    const int ox = px / b, oy = py / b;
    const int64_t total = n_panels * ox * oy * 2;
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
        const int c = (int)(t & 1);
        const int64_t q = t >> 1;
        const int j = (int)(q % oy);
        const int i = (int)((q / oy) % ox);
        const int64_t l = q / ((int64_t)ox * oy);
        auto at = [&](int a, int e) -> int64_t {
            return in[(((l * px) + (int64_t)(b * i + a)) * py + (b * j + e)) * 2 + c];
        };
        int64_t m = at(0, 0);
        m = at(1, 0) < m ? at(1, 0) : m;
        m = at(0, 1) < m ? at(0, 1) : m;
        m = at(1, 1) < m ? at(1, 1) : m;
        // floor division as numpy's // does
        int64_t qd = m / b;
        if ((m % b != 0) && (m < 0)) --qd;
        out[t] = qd;
    }
}

}  // namespace

extern "C" int cmtip_bin_images(const float* in, int64_t n_stack, int nx, int ny, int bin_factor, float* out,
                                void* stream) {
    if (n_stack < 0 || nx < 1 || ny < 1 || bin_factor < 1 || bin_factor > 128 || nx % bin_factor ||
        ny % bin_factor || (n_stack > 0 && (!in || !out))) {
        set_error("bin_images: bad arguments (the panel shape must be divisible by bin_factor <= 128)");
        return CMTIP_ERR_ARG;
    }
    const int64_t total = n_stack * (nx / bin_factor) * (int64_t)(ny / bin_factor);
    if (total == 0) return CMTIP_OK;
    hipLaunchKernelGGL(bin_images_kernel, dim3(grid_for(total)), dim3(kT), 0, as_stream(stream), in, n_stack, nx, ny,
                       bin_factor, out);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_resolution_select(const float* s, int64_t stride, int64_t P, double res_limit, uint8_t* flags,
                                       int64_t* idx, int64_t* count, void* stream) {
    if (P < 0 || stride < P || (P > 0 && (!s || !flags || !idx || !count))) {
        set_error("resolution_select: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (P == 0) return CMTIP_OK;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(resolution_flags_kernel, dim3(grid_for(P)), dim3(kT), 0, st, s, stride, P, (float)res_limit,
                       flags);
    hipLaunchKernelGGL(compact_kernel, dim3(1), dim3(kScan), 0, st, (const uint8_t*)flags, P, idx, count);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_gather_columns(const void* in, int elem_bytes, int64_t rows, int64_t P_in, const int64_t* idx,
                                    int64_t P_out, void* out, void* stream) {
    if (rows < 0 || P_in < 0 || P_out < 0 || (elem_bytes != 4 && elem_bytes != 8) ||
        (rows > 0 && P_out > 0 && (!in || !idx || !out))) {
        set_error("gather_columns: bad arguments");
        return CMTIP_ERR_ARG;
    }
    const int64_t total = rows * P_out;
    if (total == 0) return CMTIP_OK;
    hipStream_t st = as_stream(stream);
    if (elem_bytes == 4)
        hipLaunchKernelGGL(gather_cols_kernel<float>, dim3(grid_for(total)), dim3(kT), 0, st, (const float*)in, rows,
                           P_in, idx, P_out, (float*)out);
    else
        hipLaunchKernelGGL(gather_cols_kernel<double>, dim3(grid_for(total)), dim3(kT), 0, st, (const double*)in,
                           rows, P_in, idx, P_out, (double*)out);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_assemble_index(const int64_t* index_map, int64_t n_pixels, int64_t X, int64_t Y, int64_t* src,
                                    void* stream) {
    if (n_pixels < 0 || X < 1 || Y < 1 || !src || (n_pixels > 0 && !index_map)) {
        set_error("assemble_index: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipStream_t st = as_stream(stream);
    const int64_t n = X * Y;
    unsigned long long* s = reinterpret_cast<unsigned long long*>(src);
    hipLaunchKernelGGL(fill_i64_kernel, dim3(grid_for(n)), dim3(kT), 0, st, src, n, (int64_t)0);
    if (n_pixels > 0)
        hipLaunchKernelGGL(assemble_src_kernel, dim3(grid_for(n_pixels)), dim3(kT), 0, st, index_map, n_pixels, X, Y,
                           s);
    hipLaunchKernelGGL(src_to_index_kernel, dim3(grid_for(n)), dim3(kT), 0, st, (const unsigned long long*)s, n, src);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_bin_index_map(const int64_t* in, int64_t n_panels, int px, int py, int bin_factor, int64_t* out,
                                   void* stream) {
    if (n_panels < 0 || px < 2 || py < 2 || bin_factor < 2 || px % bin_factor || py % bin_factor ||
        (n_panels > 0 && (!in || !out))) {
        set_error("bin_index_map: bad arguments (panel shape divisible by bin_factor >= 2)");
        return CMTIP_ERR_ARG;
    }
    const int64_t total = n_panels * (px / bin_factor) * (int64_t)(py / bin_factor) * 2;
    if (total == 0) return CMTIP_OK;
    hipLaunchKernelGGL(bin_index_map_kernel, dim3(grid_for(total)), dim3(kT), 0, as_stream(stream), in, n_panels, px,
                       py, bin_factor, out);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}
