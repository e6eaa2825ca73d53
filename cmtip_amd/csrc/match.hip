// Orientation matching: weighted Euclidean distance between model slices and images,
// reduced to a per-image argmin (FP32) or top-k followed by FP32 rescoring (BF16).
//
// Reference: cmtip/alignment/align.py:113-131 (per-batch std scaling, weights,
// sklearn euclidean_distances, np.argmin over references).  The distance matrix is
// never materialised: the cross term <a_r, w^2 b_i> runs on MFMA and the epilogue
// forms d2 = gamma^2 an[r] - 2 gamma cross + bn[i] and reduces over r in registers,
// then with one 64-bit atomicMin per (image, workgroup column).  Keys pack
// (float bits of d2 >= 0) << 32 | r, so the unsigned order is (d2, r) and ties keep
// the smallest reference index, as np.argmin does.
//
// The BF16 form can also take its cross term from the hipBLASLt GEMM
// (cmtip_match_cross_bf16, a plain library GEMM on MFMA with FP32 accumulation and
// output) and reduce the FP32 cross-term block to per-image candidates in a separate
// pass (cmtip_match_topk_cross).
#include <hipblaslt/hipblaslt.h>

#include <stdlib.h>

#include <map>
#include <mutex>
#include <tuple>
#include <utility>

#include "common.h"

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kThreads = 256;
constexpr int BM = 128;  // references per workgroup tile
constexpr int BN = 128;  // images per workgroup tile
constexpr unsigned long long kEmpty = ~0ull;

// host: null or 16-byte aligned (the float4 forms of the row kernels)
inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

__device__ __forceinline__ unsigned long long pack_key(float d2, uint32_t r) {
    d2 = fmaxf(d2, 0.0f);  // sklearn clamps negative squared distances to 0
    return ((unsigned long long)__float_as_uint(d2) << 32) | (unsigned long long)r;
}

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
    const uint32_t lo = __shfl_xor((uint32_t)v, m, 64);
    const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ unsigned long long umin64(unsigned long long a, unsigned long long b) {
    return a < b ? a : b;
}

// XCD-aware tile order.  Workgroups are dealt to the 8 XCDs in turn (workgroup b runs
// on XCD b % 8), and each XCD has its own L2.  Workgroup b is given position
// v = (b % 8) * (n8 / 8) + b / 8 of a grouped tile order (n8: the tile count rounded
// down to a multiple of 8; the remainder keeps v = b), so the workgroups resident on one
// XCD at a time hold consecutive positions; the grouped order walks kGroupM tile rows
// before the next tile column, so those positions form a compact kGroupM-row block of
// tiles whose A and B panels the XCD's L2 shares.  (The BF16 direct-to-LDS kernel walks
// 4-row groups: with its 256 x 256 tiles, 4 x 8 blocks measured 3 % faster than 8 x 4.)
constexpr int kGroupM = 8;
template <int G = kGroupM>
__device__ __forceinline__ void tile_of(int b, int tiles_m, int tiles, int* tm, int* tn) {
    const int n8 = tiles & ~7;
    const int v = b < n8 ? (b & 7) * (n8 >> 3) + (b >> 3) : b;
    const int per_group = G * (tiles / tiles_m);  // tiles_n columns of G rows
    const int g = v / per_group;
    const int first = g * G;
    const int gm = min(tiles_m - first, G);
    const int r = v - g * per_group;
    *tm = first + r % gm;
    *tn = r / gm;
}

// ---------------------------------------------------------------- small reductions

__global__ void weighted_norms_kernel(const float* __restrict__ X, int64_t P, const float* __restrict__ w2,
                                      float* __restrict__ out) {
    __shared__ double red[kThreads / 64];
    const float* row = X + (int64_t)blockIdx.x * P;
    double s = 0.0;
    for (int64_t p = threadIdx.x; p < P; p += kThreads) {
        const double v = (double)row[p];
        s += (double)w2[p] * v * v;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kThreads / 64; ++i) t += red[i];
        out[blockIdx.x] = (float)t;
    }
}

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__global__ void prepare_images_kernel(const float* __restrict__ B, int64_t total, int64_t P,
                                      const float* __restrict__ w2, const float* __restrict__ c,
                                      float* __restrict__ Bw, uint16_t* __restrict__ Bw_bf16) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= total) return;
    const int64_t p = i % P;
    const float v = w2[p] * (c ? B[i] - c[p] : B[i]);
    if (Bw) Bw[i] = v;
    if (Bw_bf16) Bw_bf16[i] = f32_to_bf16(v);
}

// The same per row (one workgroup per image) for P a multiple of 4 and 16-byte aligned
// rows: float4 loads, one 8-byte store of four bfloat16 values.  With norms, also
// norms[row] = sum w2 (b - c)^2 in FP64, the same values and summation order as
// center_rows4_kernel with gamma = 1 (one pass over the images instead of two).
__global__ void __launch_bounds__(kThreads) prepare_images4_kernel(const float* __restrict__ B, int64_t P,
                                                                   const float* __restrict__ w2,
                                                                   const float* __restrict__ c,
                                                                   float* __restrict__ Bw,
                                                                   uint16_t* __restrict__ Bw_bf16,
                                                                   float* __restrict__ norms) {
    __shared__ double red[kThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * P, P4 = P >> 2;
    const float4* B4 = reinterpret_cast<const float4*>(B + base);
    const float4* w4 = reinterpret_cast<const float4*>(w2);
    const float4* c4 = reinterpret_cast<const float4*>(c);
    double s = 0.0;
#pragma unroll 4
    for (int64_t j = threadIdx.x; j < P4; j += kThreads) {
        const float4 b = B4[j], w = w4[j];
        const float4 cc = c ? c4[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 y = make_float4(b.x - cc.x, b.y - cc.y, b.z - cc.z, b.w - cc.w);
        const float4 v = make_float4(w.x * y.x, w.y * y.y, w.z * y.z, w.w * y.w);
        if (norms)
            s += ((double)w.x * (double)y.x * (double)y.x + (double)w.y * (double)y.y * (double)y.y) +
                 ((double)w.z * (double)y.z * (double)y.z + (double)w.w * (double)y.w * (double)y.w);
        if (Bw) reinterpret_cast<float4*>(Bw + base)[j] = v;
        if (Bw_bf16)
            reinterpret_cast<uint2*>(Bw_bf16 + base)[j] =
                make_uint2((uint32_t)f32_to_bf16(v.x) | ((uint32_t)f32_to_bf16(v.y) << 16),
                           (uint32_t)f32_to_bf16(v.z) | ((uint32_t)f32_to_bf16(v.w) << 16));
    }
    if (!norms) return;  // uniform over the workgroup
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kThreads / 64; ++i) t += red[i];
        norms[blockIdx.x] = (float)t;
    }
}

// Column sums of X (rows x P) in FP64: block (pixel group of 64, kStatWaves waves)
// splits the rows over its waves; the per-wave partials are combined in LDS in a
// fixed order, so the result does not depend on scheduling.  colsum[p] = sum_r X[r,p],
// colsum[P + p] = sum_r X[r,p]^2.
constexpr int kStatWaves = 16;

__global__ void __launch_bounds__(64 * kStatWaves) column_sums_kernel(const float* __restrict__ X, int64_t rows,
                                                                      int64_t P, double* __restrict__ colsum) {
    __shared__ double red[2][kStatWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t p = (int64_t)blockIdx.x * 64 + lane;
    double s1 = 0.0, s2 = 0.0;
    if (p < P) {
        int64_t r = wave;
        for (; r + 3 * kStatWaves < rows; r += 4 * kStatWaves) {
            const float a = X[r * P + p], b = X[(r + kStatWaves) * P + p];
            const float c = X[(r + 2 * kStatWaves) * P + p], d = X[(r + 3 * kStatWaves) * P + p];
            s1 += ((double)a + (double)b) + ((double)c + (double)d);
            s2 += ((double)a * a + (double)b * b) + ((double)c * c + (double)d * d);
        }
        for (; r < rows; r += kStatWaves) {
            const double v = (double)X[r * P + p];
            s1 += v;
            s2 += v * v;
        }
    }
    red[0][wave][lane] = s1;
    red[1][wave][lane] = s2;
    __syncthreads();
    if (wave == 0 && p < P) {
        double t1 = 0.0, t2 = 0.0;
        for (int w = 0; w < kStatWaves; ++w) {
            t1 += red[0][w][lane];
            t2 += red[1][w][lane];
        }
        colsum[p] = t1;
        colsum[P + p] = t2;
    }
}

// c = colsum / rows (float32) and mom = (sum of all, sum of squares) in a fixed order
constexpr int kFinishThreads = 1024;

__global__ void __launch_bounds__(kFinishThreads) finish_stats_kernel(const double* __restrict__ colsum,
                                                                      int64_t rows, int64_t P, float* __restrict__ c,
                                                                      double* __restrict__ mom) {
    __shared__ double red[2][kFinishThreads];
    double s1 = 0.0, s2 = 0.0;
    const double inv = 1.0 / (double)(rows > 0 ? rows : 1);
    for (int64_t p = threadIdx.x; p < P; p += kFinishThreads) {
        const double a = colsum[p];
        s1 += a;
        s2 += colsum[P + p];
        if (c) c[p] = (float)(a * inv);
    }
    red[0][threadIdx.x] = s1;
    red[1][threadIdx.x] = s2;
    __syncthreads();
    for (int h = kFinishThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            red[0][threadIdx.x] += red[0][threadIdx.x + h];
            red[1][threadIdx.x] += red[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && mom) {
        mom[0] = red[0][0];
        mom[1] = red[1][0];
    }
}

// Per row: y = gamma x - c; norms[row] = sum w2 y^2 (FP64 accumulation); y is written
// as float32 (Y, may alias X) and/or bfloat16 (Y16) when those are given.
__global__ void center_rows_kernel(const float* X, int64_t P, const double* __restrict__ gamma,
                                   const float* __restrict__ c, const float* __restrict__ w2, float* Y,
                                   uint16_t* __restrict__ Y16, float* __restrict__ norms) {
    __shared__ double red[kThreads / 64];
    const float g = gamma ? (float)gamma[0] : 1.0f;
    const int64_t base = (int64_t)blockIdx.x * P;
    double s = 0.0;
    for (int64_t p = threadIdx.x; p < P; p += kThreads) {
        const float y = c ? fmaf(g, X[base + p], -c[p]) : g * X[base + p];
        s += (double)w2[p] * (double)y * (double)y;
        if (Y) Y[base + p] = y;
        if (Y16) Y16[base + p] = f32_to_bf16(y);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kThreads / 64; ++i) t += red[i];
        norms[blockIdx.x] = (float)t;
    }
}

// center_rows_kernel for P a multiple of 4 and 16-byte aligned rows (Y16 rows 8-byte
// aligned): float4 loads, float4 / 8-byte bfloat16 stores; y and its roundings are the
// same values (fmaf(g, x, -0) = g x), only the FP64 norm adds in another order.
__global__ void __launch_bounds__(kThreads) center_rows4_kernel(const float* X, int64_t P,
                                                                const double* __restrict__ gamma,
                                                                const float* __restrict__ c,
                                                                const float* __restrict__ w2, float* Y,
                                                                uint16_t* __restrict__ Y16,
                                                                float* __restrict__ norms) {
    __shared__ double red[kThreads / 64];
    const float g = gamma ? (float)gamma[0] : 1.0f;
    const int64_t base = (int64_t)blockIdx.x * P, P4 = P >> 2;
    const float4* X4 = reinterpret_cast<const float4*>(X + base);
    const float4* c4 = reinterpret_cast<const float4*>(c);
    const float4* w4 = reinterpret_cast<const float4*>(w2);
    double s = 0.0;
#pragma unroll 4
    for (int64_t j = threadIdx.x; j < P4; j += kThreads) {
        const float4 x = X4[j], w = w4[j];
        const float4 cc = c ? c4[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 y = make_float4(fmaf(g, x.x, -cc.x), fmaf(g, x.y, -cc.y), fmaf(g, x.z, -cc.z),
                                     fmaf(g, x.w, -cc.w));
        s += ((double)w.x * (double)y.x * (double)y.x + (double)w.y * (double)y.y * (double)y.y) +
             ((double)w.z * (double)y.z * (double)y.z + (double)w.w * (double)y.w * (double)y.w);
        if (Y) reinterpret_cast<float4*>(Y + base)[j] = y;
        if (Y16)
            reinterpret_cast<uint2*>(Y16 + base)[j] =
                make_uint2((uint32_t)f32_to_bf16(y.x) | ((uint32_t)f32_to_bf16(y.y) << 16),
                           (uint32_t)f32_to_bf16(y.z) | ((uint32_t)f32_to_bf16(y.w) << 16));
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kThreads / 64; ++i) t += red[i];
        norms[blockIdx.x] = (float)t;
    }
}

// gamma[0..n) = float32(std_img / std_model); one thread computes, the block writes
__global__ void match_gamma_kernel(const double* __restrict__ img_mom, double img_count,
                                   const double* __restrict__ model_mom, double model_count, int64_t n,
                                   double* __restrict__ gamma) {
#pragma clang fp contract(off)
    __shared__ double g;
    if (threadIdx.x == 0) {
        const double ci = img_count > 0.0 ? img_count : fmax(img_mom[2], 1.0);
        const double mi = img_mom[0] / ci;
        const double si = sqrt(fmax(img_mom[1] / ci - mi * mi, 0.0));
        const double mm = model_mom[0] / model_count;
        const double sm = sqrt(fmax(model_mom[1] / model_count - mm * mm, 0.0));
        g = (double)(float)(si / sm);
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < n; i += kThreads) gamma[i] = g;
}

__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, int64_t count, unsigned long long v) {
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < count; i += stride) p[i] = v;
}

// ---------------------------------------------------------------- per-image pixel masks
//
// With a mask m_i (1 = valid pixel) the distance is
//     d2[r,i] = sum_p m_ip w_p^2 ((gamma a_rp - c_p) - (b_ip - c_p))^2
//             = sum_p [y_rp^2, y_rp] . [m_ip w_p^2, -2 m_ip w_p^2 (b_ip - c_p)] + bn_i,  y = gamma a - c
// i.e. one cross term over 2P columns: the model rows become [y^2, y] and the image
// rows [-m w^2 / 2, m w^2 (b - c)], so the matching kernels run unchanged with an = 0
// (they form -2 <a', b'> + bn).

// colsum[p] = sum_r m x, colsum[P + p] = sum_r m x^2, colsum[2P + p] = sum_r m (FP64, fixed order)
__global__ void __launch_bounds__(64 * kStatWaves) column_sums_masked_kernel(const float* __restrict__ X,
                                                                             const uint8_t* __restrict__ mask,
                                                                             int64_t rows, int64_t P,
                                                                             double* __restrict__ colsum) {
    __shared__ double red[3][kStatWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t p = (int64_t)blockIdx.x * 64 + lane;
    double s1 = 0.0, s2 = 0.0, s0 = 0.0;
    if (p < P) {
        for (int64_t r = wave; r < rows; r += kStatWaves) {
            const int64_t o = r * P + p;
            if (mask[o]) {
                const double v = (double)X[o];
                s1 += v;
                s2 += v * v;
                s0 += 1.0;
            }
        }
    }
    red[0][wave][lane] = s1;
    red[1][wave][lane] = s2;
    red[2][wave][lane] = s0;
    __syncthreads();
    if (wave == 0 && p < P) {
        double t1 = 0.0, t2 = 0.0, t0 = 0.0;
        for (int w = 0; w < kStatWaves; ++w) {
            t1 += red[0][w][lane];
            t2 += red[1][w][lane];
            t0 += red[2][w][lane];
        }
        colsum[p] = t1;
        colsum[P + p] = t2;
        colsum[2 * P + p] = t0;
    }
}

// c = (sum m x) / max(sum m, 1); mom = (sum m x, sum m x^2, sum m) over everything
__global__ void finish_stats_masked_kernel(const double* __restrict__ colsum, int64_t P, float* __restrict__ c,
                                           double* __restrict__ mom) {
    __shared__ double red[3][kThreads];
    double s1 = 0.0, s2 = 0.0, s0 = 0.0;
    for (int64_t p = threadIdx.x; p < P; p += kThreads) {
        const double a = colsum[p], cnt = colsum[2 * P + p];
        s1 += a;
        s2 += colsum[P + p];
        s0 += cnt;
        if (c) c[p] = (float)(a / (cnt > 0.0 ? cnt : 1.0));
    }
    red[0][threadIdx.x] = s1;
    red[1][threadIdx.x] = s2;
    red[2][threadIdx.x] = s0;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0 && mom) {
        mom[0] = red[0][0];
        mom[1] = red[1][0];
        mom[2] = red[2][0];
    }
}

// image rows -> [-m w^2 / 2, m w^2 (b - c)] (2P columns, float32 and/or bf16); bn = sum m w^2 (b - c)^2
__global__ void prepare_masked_kernel(const float* __restrict__ B, const uint8_t* __restrict__ mask, int64_t P,
                                      const float* __restrict__ w2, const float* __restrict__ c,
                                      float* __restrict__ Bw, uint16_t* __restrict__ Bw16, float* __restrict__ bn) {
    __shared__ double red[kThreads / 64];
    const int64_t row = blockIdx.x;
    double s = 0.0;
    for (int64_t p = threadIdx.x; p < P; p += kThreads) {
        const int64_t o = row * P + p;
        const float m = mask[o] ? w2[p] : 0.0f;
        const float d = B[o] - c[p];
        const float lo = -0.5f * m, hi = m * d;
        s += (double)m * (double)d * (double)d;
        if (Bw) {  // [-m w^2 / 2, m w^2 (b - c)]
            const int64_t q = row * 2 * P;
            Bw[q + p] = lo;
            Bw[q + P + p] = hi;
        }
        if (Bw16) {  // [-m w^2 / 2, -m w^2 / 2, m w^2 (b - c)]: pairs with the split y^2
            const int64_t q = row * 3 * P;
            const uint16_t l16 = f32_to_bf16(lo);
            Bw16[q + p] = l16;
            Bw16[q + P + p] = l16;
            Bw16[q + 2 * P + p] = f32_to_bf16(hi);
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kThreads / 64; ++i) t += red[i];
        bn[row] = (float)t;
    }
}

// model rows -> [y^2, y] with y = gamma a - c (2P float32 columns) and/or
// [hi(y^2), lo(y^2), y] (3P bfloat16 columns): y^2 carried as the sum of two bfloat16
// values (hi = bf16(y^2), lo = bf16(y^2 - hi)), which keeps the masked squared term to
// ~16 significant bits; with beamstop masks the best-to-second distance gaps are small
// enough that a single bfloat16 y^2 misplaces the best reference outside the top 8
__global__ void center_rows_sq_kernel(const float* __restrict__ X, int64_t P, const double* __restrict__ gamma,
                                      const float* __restrict__ c, float* __restrict__ Y, uint16_t* __restrict__ Y16) {
    const float g = gamma ? (float)gamma[0] : 1.0f;
    const int64_t row = blockIdx.x;
    for (int64_t p = threadIdx.x; p < P; p += kThreads) {
        const float y = fmaf(g, X[row * P + p], -c[p]);
        const float y2 = y * y;
        if (Y) {
            const int64_t q = row * 2 * P;
            Y[q + p] = y2;
            Y[q + P + p] = y;
        }
        if (Y16) {
            const int64_t q = row * 3 * P;
            const uint16_t h = f32_to_bf16(y2);
            Y16[q + p] = h;
            Y16[q + P + p] = f32_to_bf16(y2 - __uint_as_float((uint32_t)h << 16));
            Y16[q + 2 * P + p] = f32_to_bf16(y);
        }
    }
}

// ---------------------------------------------------------------- FP32 MFMA argmin
//
// WM x WN waves; wave (wm, wn) owns TM x TN tiles of v_mfma_f32_16x16x4f32 (16 references
// x 16 images each), so a workgroup covers BMt = 16 WM TM references by BNt = 16 WN TN
// images.  Operand layouts of that instruction: A[i][k] in lane i + 16 k, B[k][j] in lane
// j + 16 k, and D[i][j] in lane j + 16 (i / 4), register i % 4.  K is staged through
// double-buffered LDS in slabs of BK = 32 columns, each loaded from global memory two
// slabs ahead (two register sets).  The order of the K terms inside a slab is free (the
// same for A and B): lane group lk (lanes 16 lk .. 16 lk + 15) takes slab columns
// 8 lk .. 8 lk + 7 in its eight MFMA steps, so a lane reads its operands as two 16-byte
// LDS reads per tile row per slab instead of eight 4-byte reads.  The LDS row pitch of
// 36 floats keeps those reads 16-byte aligned, and 8 rows of one lane group start on 8
// distinct 4-bank groups.
//
// The launch uses 128 x 128 tiles (four waves of 64 x 64, two workgroups per CU).  At
// 10000 references x 5000 images its 3160 tiles take 6.17 rounds of 512 resident
// workgroups; the 88 tiles of the seventh round are split along K (SplitTail below) so
// that round is not mostly idle.  A 256 x 160 tile (eight waves of 64 x 80, one
// workgroup per CU: exactly 5 rounds) measured slower, 16.7 against 13.7 ms: with one
// workgroup per CU nothing covers its barrier waits.

template <int WM, int WN, int TM, int TN, int BK>
struct F32Tile {
    static constexpr int kWaves = WM * WN;
    static constexpr int kThreads = 64 * kWaves;
    static constexpr int BMt = 16 * WM * TM;
    static constexpr int BNt = 16 * WN * TN;
    static constexpr int LD = BK + 4;                                   // a multiple of 4 floats
    static constexpr int TPR = BK / 4;                                  // threads (float4 each) per row
    static constexpr int RPP = kThreads / TPR;                          // rows per staging pass
    static constexpr int NA = (BMt + RPP - 1) / RPP;                    // float4 per thread, A
    static constexpr int NB = (BNt + RPP - 1) / RPP;                    // float4 per thread, B
    static constexpr int kWgPerCu = kWaves == 4 ? 2 : 1;
    static_assert(kThreads % TPR == 0, "whole rows per staging pass");
};

// float4 u of this thread for the slab at column k0 of an operand of `rows` rows staged
// from row0 (NR float4 per thread); rows past the tile (the last pass of a tile whose
// rows do not divide evenly) are skipped, rows past the end of the operand read zeros, or
// with FULL (a whole slab of 16-byte aligned rows) are clamped to the last row (their
// products land in rows / columns the epilogue skips), so those loads carry no branches.
template <class T, int NR, int TILE_ROWS, bool FULL>
__device__ __forceinline__ void load_slab_f32(const float* __restrict__ X, int64_t rows, int64_t P,
                                              int64_t row0, int64_t k0, float4 (&reg)[NR]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int lr = t / T::TPR + T::RPP * i;
        if (TILE_ROWS % T::RPP != 0 && i == NR - 1 && lr >= TILE_ROWS) continue;
        int64_t r = row0 + lr;
        const int64_t k = k0 + (t % T::TPR) * 4;
        if (FULL) {
            r = r < rows ? r : rows - 1;
            reg[i] = *reinterpret_cast<const float4*>(X + r * P + k);
            continue;
        }
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (r < rows) {
            const float* src = X + r * P + k;
            if (k + 3 < P && ((reinterpret_cast<uintptr_t>(src) & 15) == 0)) {
                v = *reinterpret_cast<const float4*>(src);
            } else {
                if (k < P) v.x = src[0];
                if (k + 1 < P) v.y = src[1];
                if (k + 2 < P) v.z = src[2];
                if (k + 3 < P) v.w = src[3];
            }
        }
        reg[i] = v;
    }
}

template <class T, int NR, int TILE_ROWS>
__device__ __forceinline__ void store_slab_f32(float* __restrict__ S_, const float4 (&reg)[NR]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int lr = t / T::TPR + T::RPP * i;
        if (TILE_ROWS % T::RPP != 0 && i == NR - 1 && lr >= TILE_ROWS) continue;
        *reinterpret_cast<float4*>(S_ + lr * T::LD + (t % T::TPR) * 4) = reg[i];
    }
}

__device__ __forceinline__ float f4_at(const float4& v, int c) {
    return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w));
}

// WRITE_DIST: instead of the argmin, write the full matrix D[r, i] = sqrt(max(d2, 0))
// (float32, row-major n_ref x n_img) as sklearn's euclidean_distances returns it.
// Split tail: the last `rem` tiles of the launch (the partial last round of resident
// workgroups) are each cut into `split` runs of consecutive K slabs, taken by workgroups
// full .. full + rem * split - 1.  A run writes its partial tile to ws (per lane, its
// accumulator registers) and counts itself in cnt[tile]; the run that completes the count
// adds the split partials in run order (the same order every launch, so the result is
// deterministic), resets the counter and finishes the tile.  split = 1: no tail split.
struct SplitTail {
    int full, rem, split;
    float* ws;
    int* cnt;
};

template <int WM, int WN, int TM, int TN, bool WRITE_DIST, bool ALIGNED>
__global__ void __launch_bounds__(64 * WM * WN, (WM * WN == 4 ? 2 : 1))
    match_argmin_f32_kernel(const float* __restrict__ A, int64_t n_ref, uint32_t ref_offset,
                            const float* __restrict__ Bw, int64_t n_img, int64_t P,
                            const float* __restrict__ an, const float* __restrict__ bn,
                            const double* __restrict__ gamma, unsigned long long* __restrict__ best,
                            float* __restrict__ dist, int tiles_m, SplitTail sp) {
    constexpr int BK = 32;
    using T = F32Tile<WM, WN, TM, TN, BK>;
    __shared__ __align__(16) float As[2][T::BMt * T::LD];
    __shared__ __align__(16) float Bs[2][T::BNt * T::LD];
    __shared__ int last_run;
    const int tiles = sp.full + sp.rem;
    int b = blockIdx.x, run = 0, tail = -1;
    if (b >= sp.full) {
        tail = (b - sp.full) / sp.split;
        run = (b - sp.full) - tail * sp.split;
        b = sp.full + tail;
    }
    int tm, tn;
    tile_of<>(b, tiles_m, tiles, &tm, &tn);
    const int64_t m0 = (int64_t)tm * T::BMt, n0 = (int64_t)tn * T::BNt;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave % WM, wn = wave / WM;
    const int li = lane & 15, lk = lane >> 4;

    floatx4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    const int nk_all = (int)((P + BK - 1) / BK);
    // this workgroup's slabs [k0, k0 + nk) (all of them unless the tile is split)
    const int k0 = tail < 0 ? 0 : (int)((int64_t)nk_all * run / sp.split);
    const int nk = tail < 0 ? nk_all : (int)((int64_t)nk_all * (run + 1) / sp.split) - k0;
    // slabs [0, nfull) are whole BK-column slabs; with ALIGNED they load without branches
    const int nfull = ALIGNED ? (int)(P / BK) : 0;
    auto load = [&](int u, float4 (&ra)[T::NA], float4 (&rb)[T::NB]) {
        const int kt = k0 + u;
        if (kt < nfull) {
            load_slab_f32<T, T::NA, T::BMt, true>(A, n_ref, P, m0, (int64_t)kt * BK, ra);
            load_slab_f32<T, T::NB, T::BNt, true>(Bw, n_img, P, n0, (int64_t)kt * BK, rb);
        } else {
            load_slab_f32<T, T::NA, T::BMt, false>(A, n_ref, P, m0, (int64_t)kt * BK, ra);
            load_slab_f32<T, T::NB, T::BNt, false>(Bw, n_img, P, n0, (int64_t)kt * BK, rb);
        }
    };
    auto store = [&](int buf, const float4 (&ra)[T::NA], const float4 (&rb)[T::NB]) {
        store_slab_f32<T, T::NA, T::BMt>(As[buf], ra);
        store_slab_f32<T, T::NB, T::BNt>(Bs[buf], rb);
    };
    auto compute = [&](int buf) {
        // a wave in its LDS-read + MFMA phase is issued first: of the two workgroups on a
        // CU, the one between barriers keeps the matrix cores fed while the other stores
        // its next slab (10000 x 5000 x 16384: 13.07 -> 12.66 ms)
        __builtin_amdgcn_s_setprio(1);
        // lane group lk takes the slab's columns lk * BK / 4 .. (lk + 1) * BK / 4 - 1
        const float* as = As[buf] + (wm * TM * 16 + li) * T::LD + lk * (BK / 4);
        const float* bs = Bs[buf] + (wn * TN * 16 + li) * T::LD + lk * (BK / 4);
#pragma unroll
        for (int h = 0; h < BK / 16; ++h) {
            float4 av[TM], bv[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) av[i] = *reinterpret_cast<const float4*>(as + i * 16 * T::LD + 4 * h);
#pragma unroll
            for (int j = 0; j < TN; ++j) bv[j] = *reinterpret_cast<const float4*>(bs + j * 16 * T::LD + 4 * h);
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(f4_at(av[i], c), f4_at(bv[j], c),
                                                                         acc[i][j], 0, 0, 0);
        }
        __builtin_amdgcn_s_setprio(0);
    };
    // two register sets: slab kt + 2 is requested while slab kt is computed, so each load
    // has two slabs of MFMA work to arrive
    float4 ra[T::NA], rb[T::NB], rc[T::NA], rd[T::NB];
    load(0, ra, rb);
    if (nk > 1) load(1, rc, rd);
    for (int kt = 0; kt < nk; kt += 2) {
        store(0, ra, rb);
        __syncthreads();
        if (kt + 2 < nk) load(kt + 2, ra, rb);
        compute(0);
        if (kt + 1 < nk) {
            store(1, rc, rd);
            __syncthreads();
            if (kt + 3 < nk) load(kt + 3, rc, rd);
            compute(1);
        }
    }

    if (tail >= 0) {
        // partial tile of a split run: lane-major registers, then count the run
        float4* wsp = reinterpret_cast<float4*>(sp.ws) + ((int64_t)(tail * sp.split + run) * T::kThreads +
                                                        threadIdx.x) * (TM * TN);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                wsp[i * TN + j] = make_float4(acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]);
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) last_run = atomicAdd(sp.cnt + tail, 1) == sp.split - 1;
        __syncthreads();
        if (!last_run) return;
        __threadfence();
        if (threadIdx.x == 0) sp.cnt[tail] = 0;  // ready for the next launch
        const float4* w0 = reinterpret_cast<const float4*>(sp.ws) + ((int64_t)(tail * sp.split) * T::kThreads +
                                                                    threadIdx.x) * (TM * TN);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
        for (int r = 0; r < sp.split; ++r) {
            const float4* wr = w0 + (int64_t)r * T::kThreads * (TM * TN);
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    const float4 v = wr[i * TN + j];
                    acc[i][j] += floatx4{v.x, v.y, v.z, v.w};
                }
        }
    }

    const float g = gamma ? (float)gamma[0] : 1.0f;
    const float g2 = g * g, m2g = -2.0f * g;
    if (WRITE_DIST) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int64_t c = n0 + wn * TN * 16 + j * 16 + li;
            if (c >= n_img) continue;
            const float bnc = bn[c];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int64_t r = m0 + wm * TM * 16 + i * 16 + lk * 4 + v;
                    if (r < n_ref) {
                        const float d2 = fmaf(g2, an[r], fmaf(m2g, acc[i][j][v], bnc));
                        dist[r * n_img + c] = sqrtf(fmaxf(d2, 0.0f));
                    }
                }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int64_t c = n0 + wn * TN * 16 + j * 16 + li;
        const float bnc = c < n_img ? bn[c] : 0.f;
        unsigned long long key = kEmpty;
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t r = m0 + wm * TM * 16 + i * 16 + lk * 4 + v;
                if (r < n_ref) {
                    const float d2 = fmaf(g2, an[r], fmaf(m2g, acc[i][j][v], bnc));
                    key = umin64(key, pack_key(d2, (uint32_t)r + ref_offset));
                }
            }
        key = umin64(key, shfl_xor_u64(key, 16));
        key = umin64(key, shfl_xor_u64(key, 32));
        if (lk == 0 && c < n_img && key != kEmpty) atomicMin(best + c, key);
    }
}

// ---------------------------------------------------------------- BF16 MFMA top-k
//
// The cross term on v_mfma_f32_16x16x32_bf16 (gfx950): A[i][k] in lane i + 16 (k / 8),
// element k % 8 of an 8-wide bf16 vector; B likewise; D as above.  The epilogue keeps
// the tile's smallest keys per image and inserts them into the image's global list with
// an atomicMin cascade: atomicMin(slot s, v) leaves the smaller value in the slot and the
// larger one moves on to slot s + 1, so slot s ends as the (s+1)-th smallest key ever
// inserted, whatever the interleaving.

constexpr int BK16 = 64;
constexpr int LD16 = BK16 + 8;  // padded row stride (bf16), keeps 16-byte alignment

constexpr int kTopK = 8;       // candidate lists of up to 8 keep 8 per image (registers)
constexpr int kTopKMax = 16;   // the most candidates per image (top-k kernels and rescoring)

template <int L>
__device__ __forceinline__ void topk_insert(unsigned long long (&lst)[L], unsigned long long v) {
    if (v >= lst[L - 1]) return;
#pragma unroll
    for (int s = 0; s < L; ++s) {
        const unsigned long long lo = umin64(lst[s], v);
        v = lst[s] < v ? v : lst[s];
        lst[s] = lo;
    }
}

__device__ __forceinline__ void global_insert(unsigned long long* slots, int k, unsigned long long v) {
    for (int s = 0; s < k; ++s) {
        const unsigned long long old = atomicMin(slots + s, v);
        if (old == kEmpty) return;
        v = old > v ? old : v;
    }
}

// ---------------------------------------------------------------- BF16 MFMA top-k, register-staged slabs
//
// The general form (any P, any alignment; the direct-to-LDS kernel below takes aligned
// rows with P a multiple of 64).  A workgroup tile of BMt references x BNt images streams
// (BMt + BNt) rows of K through LDS for 2 BMt BNt K flop, so the tile size sets the
// L2 -> CU traffic the MFMA rate demands: a 128 x 128 tile at the BF16 peak (~4070 flop
// per clock per CU) needs 64 B per clock per CU, 256 x 256 half of that.  WM x WN waves
// (two per SIMD) each own TM x TN 16 x 16 MFMA tiles; slabs of 64 columns, A rows then B
// rows in one LDS stage (pitch 72 bf16: 16-byte reads of 8 consecutive rows cover all 32
// banks), double buffered, the next slab's global loads held in registers while the
// current one is multiplied.
//
// Epilogue (both kernels): the tile's d2 go to LDS in groups of kThreads / 4 images
// (pitch BMt + 4 floats), four threads per image each scan a quarter of the references
// (interleaved, so the 32 lanes of a half-wave read 32 distinct banks) keeping their L
// smallest keys, and insert them into the image's global list with global_insert.
// Top-k of one image's row of d2 in LDS (BMT references of the tile, first reference
// index ref0): the image's four lanes (4q .. 4q + 3 of one wave, part = lane & 3) scan
// references part, part + 4, ... keeping their L smallest keys, the four lists are merged
// by k rounds of a 4-lane minimum (each key belongs to one lane), and lane 0 inserts the
// merged keys into the image's global list until one is no better than its k-th entry.
// Every lane of the wave must call this (the merge shuffles); valid = 0 scans nothing.
template <int L, int BMT>
__device__ __forceinline__ void tile_topk_insert(const float* drow, bool valid, int64_t rmax, uint32_t ref0, int k,
                                                 unsigned long long* slots) {
    const int part = threadIdx.x & 3;
    unsigned long long lst[L];
#pragma unroll
    for (int s = 0; s < L; ++s) lst[s] = kEmpty;
    if (valid)
        for (int e = 0; e < BMT / 4; ++e) {
            const int rl = 4 * e + part;
            if (rl < rmax) topk_insert(lst, pack_key(drow[rl], ref0 + (uint32_t)rl));
        }
    bool done = !valid || part != 0;
    unsigned long long kth = done ? 0ull : slots[k - 1];
    for (int s = 0; s < k; ++s) {
        unsigned long long m = umin64(lst[0], shfl_xor_u64(lst[0], 1));
        m = umin64(m, shfl_xor_u64(m, 2));
        if (lst[0] == m && m != kEmpty) {
#pragma unroll
            for (int t = 0; t < L - 1; ++t) lst[t] = lst[t + 1];
            lst[L - 1] = kEmpty;
        }
        if (!done) {
            if (m == kEmpty || m >= kth) {
                done = true;
            } else {
                global_insert(slots, k, m);
                kth = slots[k - 1];
            }
        }
    }
}

template <int WM, int WN, int TM, int TN>
struct XTile {
    static constexpr int kWaves = WM * WN;
    static constexpr int kThreads = 64 * kWaves;
    static constexpr int BMt = 16 * WM * TM;                     // references per tile
    static constexpr int BNt = 16 * WN * TN;                     // images per tile
    static constexpr int kRows = BMt + BNt;                      // staged rows per slab
    static constexpr int kChunksA = BMt * (BK16 / 8);            // 16-byte chunks of A per slab
    static constexpr int kChunks = kRows * (BK16 / 8);
    static constexpr int kPerThread = (kChunks + kThreads - 1) / kThreads;
    static constexpr int kStage = kRows * LD16;                  // bf16 per LDS stage
    static constexpr int kGroup = kThreads / 4;                  // images per epilogue pass
    static constexpr int kDPitch = BMt + 4;                      // floats per image row of d2
    static constexpr size_t kStageBytes = 2 * (size_t)kStage * sizeof(uint16_t);
    static constexpr size_t kDBytes = (size_t)kGroup * kDPitch * sizeof(float);
    static constexpr size_t kLds = kStageBytes > kDBytes ? kStageBytes : kDBytes;
    static_assert(kChunksA % kThreads == 0, "A chunks of a slab split evenly over the threads");
    static_assert(kLds <= 160 * 1024, "LDS of one workgroup");
};

// 16-byte chunk u of this thread for slab k0: rows of A first, then rows of B; rows past
// the end are clamped to the last row (the epilogue skips their products).  FULL: the
// whole slab lies inside P and rows are 16-byte aligned; otherwise columns past P read 0.
template <class T, bool FULL>
__device__ __forceinline__ void xtile_load(const uint16_t* __restrict__ A, int64_t n_ref,
                                           const uint16_t* __restrict__ B, int64_t n_img, int64_t P, int64_t m0,
                                           int64_t n0, int64_t k0, uint4 (&reg)[T::kPerThread]) {
#pragma unroll
    for (int u = 0; u < T::kPerThread; ++u) {
        const int c = threadIdx.x + u * T::kThreads;
        if (T::kChunks % T::kThreads != 0 && u == T::kPerThread - 1 && c >= T::kChunks) {
            reg[u] = make_uint4(0u, 0u, 0u, 0u);
            continue;
        }
        const int row = c >> 3, kc = c & 7;
        const bool isA = u * T::kThreads < T::kChunksA;  // known at compile time
        const uint16_t* X = isA ? A : B;
        const int64_t rows = isA ? n_ref : n_img;
        int64_t r = isA ? m0 + row : n0 + (row - T::BMt);
        r = r < rows ? r : rows - 1;
        const int64_t k = k0 + kc * 8;
        const uint16_t* src = X + r * P + k;
        if (FULL) {
            reg[u] = *reinterpret_cast<const uint4*>(src);
        } else if (k + 7 < P && ((reinterpret_cast<uintptr_t>(src) & 15) == 0)) {
            reg[u] = *reinterpret_cast<const uint4*>(src);
        } else {
            uint16_t e[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) e[q] = (k + q < P) ? src[q] : (uint16_t)0;
            reg[u] = make_uint4((uint32_t)e[0] | ((uint32_t)e[1] << 16), (uint32_t)e[2] | ((uint32_t)e[3] << 16),
                                (uint32_t)e[4] | ((uint32_t)e[5] << 16), (uint32_t)e[6] | ((uint32_t)e[7] << 16));
        }
    }
}

template <class T>
__device__ __forceinline__ void xtile_store(uint16_t* __restrict__ S, const uint4 (&reg)[T::kPerThread]) {
#pragma unroll
    for (int u = 0; u < T::kPerThread; ++u) {
        const int c = threadIdx.x + u * T::kThreads;
        if (T::kChunks % T::kThreads != 0 && u == T::kPerThread - 1 && c >= T::kChunks) continue;
        *reinterpret_cast<uint4*>(S + (c >> 3) * LD16 + (c & 7) * 8) = reg[u];
    }
}

template <int WM, int WN, int TM, int TN, int L, bool ALIGNED>
__global__ void __launch_bounds__(64 * WM * WN, 1)
    cross_topk_bf16_kernel(const uint16_t* __restrict__ A, int64_t n_ref, uint32_t ref_offset,
                           const uint16_t* __restrict__ Bw, int64_t n_img, int64_t P,
                           const float* __restrict__ an, const float* __restrict__ bn,
                           const double* __restrict__ gamma, int k, unsigned long long* __restrict__ cand,
                           int tiles_m) {
    using T = XTile<WM, WN, TM, TN>;
    extern __shared__ __attribute__((aligned(16))) unsigned char xlds[];
    uint16_t* S0 = reinterpret_cast<uint16_t*>(xlds);
    int tm, tn;
    tile_of<>(blockIdx.x, tiles_m, gridDim.x, &tm, &tn);
    const int64_t m0 = (int64_t)tm * T::BMt, n0 = (int64_t)tn * T::BNt;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave % WM, wn = wave / WM;
    const int li = lane & 15, lk = lane >> 4;

    floatx4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    const int nk = (int)((P + BK16 - 1) / BK16);
    const int nfull = ALIGNED ? (int)(P / BK16) : 0;
    uint4 ra[T::kPerThread];
    auto load = [&](int kt, uint4 (&reg)[T::kPerThread]) {
        if (kt < nfull)
            xtile_load<T, true>(A, n_ref, Bw, n_img, P, m0, n0, (int64_t)kt * BK16, reg);
        else
            xtile_load<T, false>(A, n_ref, Bw, n_img, P, m0, n0, (int64_t)kt * BK16, reg);
    };
    auto compute = [&](int kt) {
        const uint16_t* S = S0 + (kt & 1) * T::kStage;
        const uint16_t* as = S + (wm * TM * 16 + li) * LD16 + lk * 8;
        const uint16_t* bs = S + (T::BMt + wn * TN * 16 + li) * LD16 + lk * 8;
#pragma unroll
        for (int kk = 0; kk < BK16 / 32; ++kk) {
            bf16x8 af[TM], bf[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) af[i] = *reinterpret_cast<const bf16x8*>(as + i * 16 * LD16 + kk * 32);
#pragma unroll
            for (int j = 0; j < TN; ++j) bf[j] = *reinterpret_cast<const bf16x8*>(bs + j * 16 * LD16 + kk * 32);
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    };
    // slab kt + 1 is loaded into registers while slab kt is multiplied
    load(0, ra);
    xtile_store<T>(S0, ra);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        if (kt + 1 < nk) load(kt + 1, ra);
        compute(kt);
        if (kt + 1 < nk) xtile_store<T>(S0 + ((kt + 1) & 1) * T::kStage, ra);
        __syncthreads();
    }

    // d2 = gamma^2 an[r] - 2 gamma cross + bn[c] into LDS, image-major, 128 images a pass
    float* D = reinterpret_cast<float*>(xlds);
    const float g = gamma ? (float)gamma[0] : 1.0f;
    const float g2 = g * g, m2g = -2.0f * g;
    float anr[TM][4];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t r = m0 + wm * TM * 16 + i * 16 + lk * 4 + v;
            anr[i][v] = r < n_ref ? an[r] : 0.f;
        }
    for (int g0 = 0; g0 < T::BNt; g0 += T::kGroup) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int cl = wn * TN * 16 + j * 16 + li;
            const int q = cl - g0;
            if (q < 0 || q >= T::kGroup) continue;
            const int64_t c = n0 + cl;
            const float bnc = c < n_img ? bn[c] : 0.f;
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    D[q * T::kDPitch + wm * TM * 16 + i * 16 + lk * 4 + v] =
                        fmaf(g2, anr[i][v], fmaf(m2g, acc[i][j][v], bnc));
        }
        __syncthreads();
        const int q = threadIdx.x >> 2;
        const int64_t c = n0 + g0 + q;
        tile_topk_insert<L, T::BMt>(D + q * T::kDPitch, g0 + q < T::BNt && c < n_img, n_ref - m0,
                                    (uint32_t)m0 + ref_offset, k, cand + c * k);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- BF16 top-k, direct-to-LDS pipeline
//
// The default BF16 kernel for aligned rows with P a multiple of 64: 256 references x 256
// images per workgroup, four waves (one per SIMD) each owning 128 x 128 (8 x 8 MFMA
// tiles, 256 accumulators in AGPRs), which halves the LDS operand reads per flop of a
// 128 x 64 wave tile (16 16-byte reads per 64 MFMAs).  Operand slabs of 32 columns are
// fetched by global_load_lds_dwordx4, which writes each lane's 16 bytes straight to LDS
// (a wave instruction fills 1 KB: 16 rows of 64 bytes), so no registers stage them; S = 4
// stages of 32 KB.  Step kt: wait until this wave's loads of slab kt + 1 have landed and
// its LDS reads of slab kt are done, barrier (then every wave's have), request slab
// kt + S into the stage slab kt used, read slab kt + 1's operands into the second
// register set interleaved with slab kt's 64 MFMAs from the first (one read per four
// MFMAs, fixed with sched_group_barrier).  Three slabs (96 KB per CU) stay in flight.
// Rows are 64 bytes (four 16-byte chunks); chunk c of row r is stored at position
// c ^ f(r), f(r) = ((r >> 1) + (r >> 3)) & 3, so the 16-byte operand reads of 8 (or 16)
// consecutive rows at one chunk hit distinct banks.
//
// Measured at 10000 x 5000 x 16384 (and 51200 x 20480 x 16384): the loop is bound by
// the global -> LDS traffic, 16.8 MB per tile: the copies alone take as long as the
// whole kernel, and they run at about 9 TB/s with the tile order above; with every tile
// reading the same L2-resident rows the same loop reaches 1.7 PFLOP/s.  The XCD-aware
// order is worth 1.7x at the larger shape (plain order 648 TFLOP/s).
constexpr int BK32 = 32;

template <int WM, int WN, int TM, int TN, int S>
struct DTile {
    static constexpr int kWaves = WM * WN;
    static constexpr int kThreads = 64 * kWaves;
    static constexpr int BMt = 16 * WM * TM;
    static constexpr int BNt = 16 * WN * TN;
    static constexpr int kRows = BMt + BNt;
    static constexpr int kBlocks = kRows / 16;                   // 1 KB wave loads per slab
    static constexpr int kMinPerWave = kBlocks / kWaves;         // loads per slab of every wave (at least)
    static constexpr int kStageBytes = kRows * BK32 * 2;
    static constexpr int kGroup = kThreads / 4;
    static constexpr int kDPitch = BMt + 4;
    static constexpr size_t kDBytes = (size_t)kGroup * kDPitch * sizeof(float);
    static constexpr size_t kLds = (size_t)S * kStageBytes > kDBytes ? (size_t)S * kStageBytes : kDBytes;
    static_assert(kRows % 16 == 0 && kMinPerWave >= 1, "whole 16-row blocks, one per wave at least");
    static_assert(S >= 3 && kMinPerWave * (S - 2) < 64, "vmcnt range");
    static_assert(kLds <= 160 * 1024, "LDS of one workgroup");
};

__device__ __forceinline__ int dswz(int r) { return ((r >> 1) + (r >> 3)) & 3; }

// Scheduling order of one pipeline step: R - 1 rounds of (Q MFMAs, one LDS read), the last
// read, then the remaining MFMAs, so the next slab's operand reads spread over this
// slab's MFMAs and the last one has MFMAs behind it before the step's barrier.
template <int Q, int... I>
__device__ __forceinline__ void sched_rounds(std::integer_sequence<int, I...>) {
    ((__builtin_amdgcn_sched_group_barrier(0x008, Q, 0), __builtin_amdgcn_sched_group_barrier(0x100, 1, 0),
      (void)I),
     ...);
}

template <int R, int Q, int REST>
__device__ __forceinline__ void sched_reads_mfma() {
    sched_rounds<Q>(std::make_integer_sequence<int, R - 1>{});
    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // the last read ahead of the last MFMAs
    __builtin_amdgcn_sched_group_barrier(0x008, Q + REST, 0);
}

// s_waitcnt with vmcnt <= N (expcnt, lgkmcnt left alone)
template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

template <int WM, int WN, int TM, int TN, int S, int L>
__global__ void __launch_bounds__(64 * WM * WN, 1)
    cross_topk_bf16_dma_kernel(const uint16_t* __restrict__ A, int64_t n_ref, uint32_t ref_offset,
                               const uint16_t* __restrict__ Bw, int64_t n_img, int64_t P,
                               const float* __restrict__ an, const float* __restrict__ bn,
                               const double* __restrict__ gamma, int k, unsigned long long* __restrict__ cand,
                               int tiles_m) {
    using T = DTile<WM, WN, TM, TN, S>;
    extern __shared__ __attribute__((aligned(16))) unsigned char xlds[];
    int tm, tn;
    tile_of<4>(blockIdx.x, tiles_m, gridDim.x, &tm, &tn);
    const int64_t m0 = (int64_t)tm * T::BMt, n0 = (int64_t)tn * T::BNt;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave % WM, wn = wave / WM;
    const int li = lane & 15, lk = lane >> 4;
    const int nk = (int)(P / BK32);

    // this lane's global row pointer and LDS block of each of its wave's loads
    constexpr int kMaxPerWave = (T::kBlocks + T::kWaves - 1) / T::kWaves;
    const uint16_t* src[kMaxPerWave];
    int nblk = 0;
#pragma unroll
    for (int u = 0; u < kMaxPerWave; ++u) {
        const int b = wave + u * T::kWaves;
        const int r = 16 * b + (lane >> 2);
        const int c = (lane & 3) ^ dswz(r);
        const bool isA = r < T::BMt;
        int64_t gr = isA ? m0 + r : n0 + (r - T::BMt);
        const int64_t rows = isA ? n_ref : n_img;
        gr = gr < rows ? gr : rows - 1;
        src[u] = (isA ? A : Bw) + gr * P + 8 * c;
        nblk += b < T::kBlocks ? 1 : 0;
    }
    auto issue = [&](int kt) {
        unsigned char* stage = xlds + (kt % S) * T::kStageBytes;
#pragma unroll
        for (int u = 0; u < kMaxPerWave; ++u) {
            if (u < nblk)
                __builtin_amdgcn_global_load_lds((const void*)(src[u] + (int64_t)kt * BK32),
                                                 (__attribute__((address_space(3))) void*)(stage +
                                                     (wave + u * T::kWaves) * 1024),
                                                 16, 0, 0);
        }
    };

    floatx4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // LDS byte offsets of this lane's operand rows within a stage
    int aoff[TM], boff[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const int r = wm * TM * 16 + i * 16 + li;
        aoff[i] = r * 64 + 16 * (lk ^ dswz(r));
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int r = T::BMt + wn * TN * 16 + j * 16 + li;
        boff[j] = r * 64 + 16 * (lk ^ dswz(r));
    }
    bf16x8 af[2][TM], bfr[2][TN];
    auto read = [&](int kt, int set) {
        const unsigned char* stage = xlds + (kt % S) * T::kStageBytes;
#pragma unroll
        for (int i = 0; i < TM; ++i) af[set][i] = *reinterpret_cast<const bf16x8*>(stage + aoff[i]);
#pragma unroll
        for (int j = 0; j < TN; ++j) bfr[set][j] = *reinterpret_cast<const bf16x8*>(stage + boff[j]);
    };
    auto mma = [&](int set) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[set][i], bfr[set][j], acc[i][j], 0, 0, 0);
    };

    // prologue: slabs 0 .. S - 1 requested, slab 0's operands read
    for (int s = 0; s < S && s < nk; ++s) issue(s);
    wait_vm<0>();
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    read(0, 0);
    // nk is even: slab kt from set 0, slab kt + 1 from set 1 (static register sets).  The
    // operand reads are unconditional (past the last slab they re-read a landed stage
    // and go unused), so no branch makes the compiler wait for them before the MFMAs.
    for (int kt = 0; kt < nk; kt += 2) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int s = kt + h;
            // this wave's loads of slab s + 1 landed (slabs s + 2 .. s + S - 1 may still be
            // in flight) and its reads of slab s done, so after the barrier slab s's stage is
            // free in every wave
            if (s + S - 1 < nk)
                wait_vm<T::kMinPerWave * (S - 2)>();
            else
                wait_vm<0>();
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            if (s + S < nk) issue(s + S);
            read(s + 1, h ^ 1);
            mma(h);
            sched_reads_mfma<TM + TN, (TM * TN) / (TM + TN), TM * TN - (TM + TN) * ((TM * TN) / (TM + TN))>();
        }
    }
    wait_vm<0>();
    __syncthreads();  // the stages are reused below

    float* D = reinterpret_cast<float*>(xlds);
    const float g = gamma ? (float)gamma[0] : 1.0f;
    const float g2 = g * g, m2g = -2.0f * g;
    for (int g0 = 0; g0 < T::BNt; g0 += T::kGroup) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int cl = wn * TN * 16 + j * 16 + li;
            const int q = cl - g0;
            if (q < 0 || q >= T::kGroup) continue;
            const int64_t c = n0 + cl;
            const float bnc = c < n_img ? bn[c] : 0.f;
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    // an[r] reloaded per image group (L1 hits) to keep registers for the lists
                    const int rl = wm * TM * 16 + i * 16 + lk * 4 + v;
                    const float anr = m0 + rl < n_ref ? an[m0 + rl] : 0.f;
                    D[q * T::kDPitch + rl] = fmaf(g2, anr, fmaf(m2g, acc[i][j][v], bnc));
                }
        }
        __syncthreads();
        const int q = threadIdx.x >> 2;
        const int64_t c = n0 + g0 + q;
        tile_topk_insert<L, T::BMt>(D + q * T::kDPitch, g0 + q < T::BNt && c < n_img, n_ref - m0,
                                    (uint32_t)m0 + ref_offset, k, cand + c * k);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- FP32 rescoring
// One workgroup per image.  rescore4_kernel (P a multiple of 4, 16-byte aligned rows)
// sweeps the pixels once in float4 quads: each quad of the image, weights and mask is
// loaded once and the same quad of all k candidate rows is loaded together (k loads in
// flight), with one FP32 accumulator per candidate; then a block reduction.
// rescore_kernel is the scalar form for other shapes: each wave takes candidates
// s = wave, wave + 4, ...

template <int L>
__global__ void __launch_bounds__(kThreads) rescore4_kernel(const float* __restrict__ A, int64_t n_ref_total,
                                                            const float* __restrict__ B, int64_t P,
                                                            const float* __restrict__ w2,
                                                            const double* __restrict__ gamma_of_ref, int k,
                                                            const unsigned long long* __restrict__ cand,
                                                            unsigned long long* __restrict__ best,
                                                            const uint8_t* __restrict__ mask) {
    __shared__ float red[kThreads / 64][L];
    const int64_t i = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float4* a4[L];
    float g[L], acc[L];
    uint32_t r_of[L];
#pragma unroll
    for (int s = 0; s < L; ++s) {
        const unsigned long long key = s < k ? cand[i * k + s] : kEmpty;
        const uint32_t r = (uint32_t)(key & 0xffffffffu);
        const bool ok = key != kEmpty && (int64_t)r < n_ref_total;
        r_of[s] = ok ? r : 0xffffffffu;  // candidates past k are never valid
        g[s] = ok ? (float)gamma_of_ref[r] : 0.f;
        a4[s] = reinterpret_cast<const float4*>(A + (ok ? (int64_t)r : 0) * P);
        acc[s] = 0.f;
    }
    const float4* b4 = reinterpret_cast<const float4*>(B + i * P);
    const float4* w4 = reinterpret_cast<const float4*>(w2);
    const uchar4* m4 = mask ? reinterpret_cast<const uchar4*>(mask + i * P) : nullptr;
    const int64_t P4 = P / 4;
    for (int64_t q = threadIdx.x; q < P4; q += kThreads) {
        const float4 bv = b4[q];
        float4 wv = w4[q];
        if (m4) {
            const uchar4 mv = m4[q];
            wv.x = mv.x ? wv.x : 0.f;
            wv.y = mv.y ? wv.y : 0.f;
            wv.z = mv.z ? wv.z : 0.f;
            wv.w = mv.w ? wv.w : 0.f;
        }
        float4 av[L];
#pragma unroll
        for (int s = 0; s < L; ++s)
            if (r_of[s] != 0xffffffffu) av[s] = a4[s][q];  // same for every thread of the block
#pragma unroll
        for (int s = 0; s < L; ++s) {
            if (r_of[s] == 0xffffffffu) continue;
            float d = fmaf(g[s], av[s].x, -bv.x);
            float t = fmaf(wv.x * d, d, acc[s]);
            d = fmaf(g[s], av[s].y, -bv.y);
            t = fmaf(wv.y * d, d, t);
            d = fmaf(g[s], av[s].z, -bv.z);
            t = fmaf(wv.z * d, d, t);
            d = fmaf(g[s], av[s].w, -bv.w);
            acc[s] = fmaf(wv.w * d, d, t);
        }
    }
#pragma unroll
    for (int s = 0; s < L; ++s) {
        float v = acc[s];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[wave][s] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = kEmpty;
        for (int s = 0; s < k; ++s) {
            if (r_of[s] == 0xffffffffu) continue;
            float v = 0.f;
            for (int w = 0; w < kThreads / 64; ++w) v += red[w][s];
            m = umin64(m, pack_key(v, r_of[s]));
        }
        best[i] = m;
    }
}


__global__ void __launch_bounds__(kThreads) rescore_kernel(const float* __restrict__ A, int64_t n_ref_total,
                                                           const float* __restrict__ B, int64_t P,
                                                           const float* __restrict__ w2,
                                                           const double* __restrict__ gamma_of_ref, int k,
                                                           const unsigned long long* __restrict__ cand,
                                                           unsigned long long* __restrict__ best,
                                                           const uint8_t* __restrict__ mask) {
    __shared__ unsigned long long red[kThreads / 64];
    const int64_t i = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float* b = B + i * P;
    unsigned long long mine = kEmpty;
    for (int s = wave; s < k; s += kThreads / 64) {
        const unsigned long long key = cand[i * k + s];
        if (key == kEmpty) continue;
        const uint32_t r = (uint32_t)(key & 0xffffffffu);
        if ((int64_t)r >= n_ref_total) continue;
        const float g = (float)gamma_of_ref[r];
        const float* a = A + (int64_t)r * P;
        float acc = 0.f;
        for (int64_t p = lane; p < P; p += 64) {
            const float dv = fmaf(g, a[p], -b[p]);
            const float wp = (mask && !mask[i * P + p]) ? 0.0f : w2[p];
            acc = fmaf(wp * dv, dv, acc);
        }
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        mine = umin64(mine, pack_key(acc, r));
    }
    if (lane == 0) red[wave] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = red[0];
        for (int q = 1; q < kThreads / 64; ++q) m = umin64(m, red[q]);
        best[i] = m;
    }
}

unsigned grid1(int64_t n, int64_t cap = 1 << 20) {
    int64_t b = (n + kThreads - 1) / kThreads;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return (unsigned)b;
}

}  // namespace

using namespace cmtip;

extern "C" int cmtip_weighted_norms(const float* X, int64_t rows, int64_t P, const float* w2, float* out,
                                    void* stream) {
    if (rows < 0 || P < 0 || (rows > 0 && (!X || !w2 || !out))) {
        set_error("weighted_norms: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (rows == 0) return CMTIP_OK;
    if (rows > 0x7fffffff) {
        set_error("weighted_norms: too many rows");
        return CMTIP_ERR_ARG;
    }
    hipLaunchKernelGGL(weighted_norms_kernel, dim3((unsigned)rows), dim3(kThreads), 0, as_stream(stream), X, P,
                       w2, out);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_image_stats(const float* X, int64_t rows, int64_t P, double* colsum, float* c, double* mom,
                                 void* stream) {
    if (rows < 0 || P < 1 || !colsum || (rows > 0 && !X)) {
        set_error("image_stats: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(column_sums_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64 * kStatWaves), 0, st, X, rows, P,
                       colsum);
    hipLaunchKernelGGL(finish_stats_kernel, dim3(1), dim3(kFinishThreads), 0, st, (const double*)colsum, rows, P, c, mom);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

// ---------------------------------------------------------------- allocation-free image statistics
//
// cmtip_image_stats with its column-sum workspace in a file-scope device array instead of
// a caller buffer, for callers that cannot allocate device memory.  In host code the
// array's name is only a stand-in for the symbol: its device address is looked up once
// per device with hipGetSymbolAddress and only that address is passed to kernels.
// Every launch overwrites the workspace entirely (column_sums_kernel writes each of its
// 2P entries), so nothing needs clearing between calls.

namespace {
constexpr int64_t kStatsScratchPixels = 262144;  // 2 x 2^18 doubles = 4 MiB per device
constexpr int kMaxDevices = 64;
__device__ __align__(256) double image_stats_scratch[2 * kStatsScratchPixels];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of image_stats_scratch, looked up on first use
// with the device current.
int stats_scratch(double** out) {
    int dev = -1;
    CMTIP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("image_stats_noalloc: device ordinal %d out of range", dev);
        return CMTIP_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        CMTIP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(image_stats_scratch)));
        if (!p) {
            set_error("image_stats_noalloc: no device address for the scratch array");
            return CMTIP_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return CMTIP_OK;
}
}  // namespace

extern "C" int cmtip_image_stats_noalloc(const float* X, int64_t rows, int64_t P, float* c, double* mom,
                                         void* stream) {
    if (rows < 0 || P < 1 || (rows > 0 && !X)) {
        set_error("image_stats_noalloc: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (P > kStatsScratchPixels) {
        set_error("image_stats_noalloc: P = %lld pixels exceeds the workspace limit of %lld", (long long)P,
                  (long long)kStatsScratchPixels);
        return CMTIP_ERR_ARG;
    }
    double* colsum = nullptr;
    CMTIP_TRY(stats_scratch(&colsum));
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(column_sums_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64 * kStatWaves), 0, st, X, rows, P,
                       colsum);
    hipLaunchKernelGGL(finish_stats_kernel, dim3(1), dim3(kFinishThreads), 0, st, (const double*)colsum, rows, P, c, mom);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_match_gamma(const double* img_mom, double img_count, const double* model_mom,
                                 double model_count, int64_t n, double* gamma, void* stream) {
    if (!img_mom || !model_mom || !(model_count > 0.0) || n < 1 || !gamma) {
        set_error("match_gamma: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipLaunchKernelGGL(match_gamma_kernel, dim3(1), dim3(kThreads), 0, as_stream(stream), img_mom, img_count,
                       model_mom, model_count, n, gamma);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_center_rows(const float* X, int64_t rows, int64_t P, const double* gamma, const float* c,
                                 const float* w2, float* Y, uint16_t* Y16, float* norms, void* stream) {
    if (rows < 0 || P < 1 || (rows > 0 && (!X || !w2 || !norms)) || rows > 0x7fffffff) {
        set_error("center_rows: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (rows == 0) return CMTIP_OK;
    const bool vec = P % 4 == 0 && aligned16(X) && aligned16(Y) && aligned16(c) && aligned16(w2) &&
                     (reinterpret_cast<uintptr_t>(Y16) & 7) == 0;
    if (vec)
        hipLaunchKernelGGL(center_rows4_kernel, dim3((unsigned)rows), dim3(kThreads), 0, as_stream(stream), X, P,
                           gamma, c, w2, Y, Y16, norms);
    else
        hipLaunchKernelGGL(center_rows_kernel, dim3((unsigned)rows), dim3(kThreads), 0, as_stream(stream), X, P,
                           gamma, c, w2, Y, Y16, norms);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_match_prepare_images(const float* B, int64_t n_img, int64_t P, const float* w2, float* Bw,
                                          uint16_t* Bw_bf16, void* stream) {
    return cmtip_match_prepare_images_centered(B, n_img, P, w2, nullptr, Bw, Bw_bf16, stream);
}

extern "C" int cmtip_match_prepare_images_centered(const float* B, int64_t n_img, int64_t P, const float* w2,
                                                   const float* c, float* Bw, uint16_t* Bw_bf16, void* stream) {
    if (n_img < 0 || P < 0 || (n_img * P > 0 && (!B || !w2 || (!Bw && !Bw_bf16)))) {
        set_error("match_prepare_images: bad arguments");
        return CMTIP_ERR_ARG;
    }
    const int64_t total = n_img * P;
    if (total == 0) return CMTIP_OK;
    if (P % 4 == 0 && n_img <= 0x7fffffffLL && aligned16(B) && aligned16(w2) && aligned16(c) && aligned16(Bw) &&
        (reinterpret_cast<uintptr_t>(Bw_bf16) & 7) == 0)
        hipLaunchKernelGGL(prepare_images4_kernel, dim3((unsigned)n_img), dim3(kThreads), 0, as_stream(stream), B, P,
                           w2, c, Bw, Bw_bf16, (float*)nullptr);
    else
        hipLaunchKernelGGL(prepare_images_kernel, dim3((unsigned)((total + kThreads - 1) / kThreads)),
                           dim3(kThreads), 0, as_stream(stream), B, total, P, w2, c, Bw, Bw_bf16);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_match_prepare_images_norms(const float* B, int64_t n_img, int64_t P, const float* w2,
                                                const float* c, float* Bw, uint16_t* Bw_bf16, float* norms,
                                                void* stream) {
    if (n_img < 0 || P < 1 || (n_img > 0 && (!B || !w2 || !norms || (!Bw && !Bw_bf16)))) {
        set_error("match_prepare_images_norms: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (n_img == 0) return CMTIP_OK;
    if (P % 4 == 0 && n_img <= 0x7fffffffLL && aligned16(B) && aligned16(w2) && aligned16(c) && aligned16(Bw) &&
        (reinterpret_cast<uintptr_t>(Bw_bf16) & 7) == 0) {
        hipLaunchKernelGGL(prepare_images4_kernel, dim3((unsigned)n_img), dim3(kThreads), 0, as_stream(stream), B, P,
                           w2, c, Bw, Bw_bf16, norms);
        CMTIP_LAUNCH_CHECK();
        return CMTIP_OK;
    }
    const int rc = cmtip_match_prepare_images_centered(B, n_img, P, w2, c, Bw, Bw_bf16, stream);
    if (rc != CMTIP_OK) return rc;
    return cmtip_center_rows(B, n_img, P, nullptr, c, w2, nullptr, nullptr, norms, stream);
}

extern "C" int cmtip_fill_u64(unsigned long long* dst, int64_t count, unsigned long long value, void* stream) {
    if (count < 0 || (count > 0 && !dst)) {
        set_error("fill_u64: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (count == 0) return CMTIP_OK;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid1(count, 65536)), dim3(kThreads), 0, as_stream(stream), dst, count,
                       value);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

static int check_match_args(int64_t n_ref, int64_t ref_offset, int64_t n_img, int64_t P, const void* A,
                            const void* B, const void* an, const void* bn, const void* out, int64_t* tiles_m,
                            int64_t* tiles) {
    if (n_ref < 0 || n_img < 0 || P < 1 || ref_offset < 0 || n_ref + ref_offset > 0xffffffffLL ||
        (n_ref > 0 && n_img > 0 && (!A || !B || !an || !bn || !out))) {
        set_error("match: bad arguments");
        return CMTIP_ERR_ARG;
    }
    *tiles_m = (n_ref + BM - 1) / BM;
    *tiles = *tiles_m * ((n_img + BN - 1) / BN);
    if (*tiles > 0x7fffffff) {
        set_error("match: problem too large for one launch");
        return CMTIP_ERR_ARG;
    }
    return CMTIP_OK;
}

// The FP32 kernel runs with 32-column slabs and two slabs of prefetch: measured at
// 10000 x 5000 x 16384 with 128 x 128 tiles, 13.98 ms against 14.21 ms with one slab of
// prefetch, and 14.69 / 14.97 ms with 16-column slabs (three workgroups per CU).
// Shapes are picked by the work the launch is paid for: rounds of resident workgroups x
// the tile area one CU holds per round (2 x 128 x 128 or 256 x 160).
// Split-K workspace of the FP32 kernel, one per (device, stream): partial tiles and
// self-resetting run counters.
struct SplitWs {
    float* ws = nullptr;
    int* cnt = nullptr;
    size_t ws_floats = 0;
};
std::mutex g_split_mu;
std::map<std::pair<int, hipStream_t>, SplitWs> g_split;
constexpr int kSplitMaxTiles = 4096;

template <int WM, int WN, int TM, int TN>
struct F32Launch {
    using T = F32Tile<WM, WN, TM, TN, 32>;
    static int64_t tiles_m(int64_t n_ref) { return (n_ref + T::BMt - 1) / T::BMt; }
    static int64_t tiles(int64_t n_ref, int64_t n_img) { return tiles_m(n_ref) * ((n_img + T::BNt - 1) / T::BNt); }
    // the partial last round of resident workgroups is split along K when it is less than
    // half full: rem tiles x split runs fill up to one round; each run keeps >= 8 slabs
    static SplitTail plan(int64_t n_ref, int64_t n_img, int64_t P, int cus) {
        const int64_t t = tiles(n_ref, n_img), slots = (int64_t)cus * T::kWgPerCu;
        const int64_t rem = t % slots, nk = (P + 31) / 32;
        SplitTail sp{(int)t, 0, 1, nullptr, nullptr};
        if (t < slots || rem == 0 || 2 * rem > slots || rem > kSplitMaxTiles) return sp;
        const int64_t split = std::min<int64_t>(std::min<int64_t>(slots / rem, 8), nk / 8);
        if (split < 2) return sp;
        sp.full = (int)(t - rem);
        sp.rem = (int)rem;
        sp.split = (int)split;
        return sp;
    }
    template <bool WD>
    static int run(bool aligned, hipStream_t st, const float* A, int64_t n_ref, uint32_t off, const float* B,
                   int64_t n_img, int64_t P, const float* an, const float* bn, const double* gamma,
                   unsigned long long* best, float* dist, int cus) {
        SplitTail sp = plan(n_ref, n_img, P, cus);
        if (sp.split > 1) {
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess) {
                set_error("match: hipGetDevice failed");
                return CMTIP_ERR_HIP;
            }
            const size_t need = (size_t)sp.rem * sp.split * T::kThreads * TM * TN * 4;
            std::lock_guard<std::mutex> lk(g_split_mu);
            SplitWs& w = g_split[std::make_pair(dev, st)];
            if (!w.cnt) {
                if (hipMalloc(&w.cnt, sizeof(int) * kSplitMaxTiles) != hipSuccess ||
                    hipMemset(w.cnt, 0, sizeof(int) * kSplitMaxTiles) != hipSuccess) {
                    set_error("match: split-K counter allocation failed");
                    return CMTIP_ERR_ALLOC;
                }
            }
            if (w.ws_floats < need) {
                if (w.ws) (void)hipFree(w.ws);
                w.ws = nullptr;
                w.ws_floats = 0;
                if (hipMalloc(&w.ws, need * sizeof(float)) != hipSuccess) {
                    set_error("match: split-K workspace allocation failed");
                    return CMTIP_ERR_ALLOC;
                }
                w.ws_floats = need;
            }
            sp.ws = w.ws;
            sp.cnt = w.cnt;
        }
        const unsigned grid = (unsigned)(sp.full + sp.rem * sp.split);
        const int tm = (int)tiles_m(n_ref);
        if (aligned)
            hipLaunchKernelGGL((match_argmin_f32_kernel<WM, WN, TM, TN, WD, true>), dim3(grid), dim3(T::kThreads), 0,
                               st, A, n_ref, off, B, n_img, P, an, bn, gamma, best, dist, tm, sp);
        else
            hipLaunchKernelGGL((match_argmin_f32_kernel<WM, WN, TM, TN, WD, false>), dim3(grid), dim3(T::kThreads),
                               0, st, A, n_ref, off, B, n_img, P, an, bn, gamma, best, dist, tm, sp);
        return CMTIP_OK;
    }
};
using F32Small = F32Launch<2, 2, 4, 4>;  // 128 x 128, two workgroups per CU

int match_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
    }
    return cus;
}

template <bool WD>
static int launch_f32(bool aligned, hipStream_t st, const float* A, int64_t n_ref, uint32_t off, const float* B,
                      int64_t n_img, int64_t P, const float* an, const float* bn, const double* gamma,
                      unsigned long long* best, float* dist) {
    return F32Small::run<WD>(aligned, st, A, n_ref, off, B, n_img, P, an, bn, gamma, best, dist, match_cus());
}

extern "C" int cmtip_match_argmin_f32(const float* A, int64_t n_ref, int64_t ref_offset, const float* Bw,
                                      int64_t n_img, int64_t P, const float* an, const float* bn,
                                      const double* gamma, unsigned long long* best, void* stream) {
    int64_t tiles_m, tiles;
    CMTIP_TRY(check_match_args(n_ref, ref_offset, n_img, P, A, Bw, an, bn, best, &tiles_m, &tiles));
    if (tiles == 0) return CMTIP_OK;
    CMTIP_TRY(launch_f32<false>(P % 4 == 0 && aligned16(A) && aligned16(Bw), as_stream(stream), A, n_ref,
                                (uint32_t)ref_offset, Bw, n_img, P, an, bn, gamma, best, (float*)nullptr));
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_eudist_f32(const float* A, int64_t n_a, const float* B, int64_t n_b, int64_t P,
                                const float* an, const float* bn, float* D, void* stream) {
    int64_t tiles_m, tiles;
    CMTIP_TRY(check_match_args(n_a, 0, n_b, P, A, B, an, bn, D, &tiles_m, &tiles));
    if (tiles == 0) return CMTIP_OK;
    CMTIP_TRY(launch_f32<true>(P % 4 == 0 && aligned16(A) && aligned16(B), as_stream(stream), A, n_a, 0u, B, n_b, P,
                               an, bn, (const double*)nullptr, (unsigned long long*)nullptr, D));
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

namespace {

// ---------------------------------------------------------------- top-k from a cross-term block
// One wave per image i: lanes scan references r = lane, lane + 64, ... of the row
// Dt[i, 0..n_ref) and keep their k smallest keys (d2, r) in registers; the image's
// earlier candidates (previous reference batches) join the lists of lanes 0..k-1; k
// rounds of a wave-wide minimum over the list heads then give the merged top-k, written
// back without atomics (ties keep the smallest reference index).
constexpr int kCrossWaves = 4;
constexpr int kCrossBatch = 4;  // loads in flight per lane before the inserts

// VEC: n_ref a multiple of 4 and 16-byte aligned Dt and an (float4 loads).  The lists
// hold the lane's smallest keys whatever the order the references arrive in, so both
// forms give the same top-k.
template <bool VEC, int L>
__global__ void __launch_bounds__(64 * kCrossWaves) topk_cross_kernel(const float* __restrict__ Dt, int64_t n_ref,
                                                                      uint32_t ref_offset, int64_t n_img,
                                                                      const float* __restrict__ an,
                                                                      const float* __restrict__ bn,
                                                                      const double* __restrict__ gamma, int k,
                                                                      unsigned long long* __restrict__ cand) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * kCrossWaves + (threadIdx.x >> 6);
    if (i >= n_img) return;  // whole waves leave together
    const float g = gamma ? (float)gamma[0] : 1.0f;
    const float g2 = g * g, m2g = -2.0f * g;
    const float bni = bn[i];
    unsigned long long lst[L];
#pragma unroll
    for (int s = 0; s < L; ++s) lst[s] = kEmpty;
    unsigned long long* slots = cand + i * k;
    if (lane < k) topk_insert(lst, slots[lane]);
    const float* row = Dt + i * n_ref;
    if (VEC) {
        const float4* row4 = reinterpret_cast<const float4*>(row);
        const float4* an4 = reinterpret_cast<const float4*>(an);
        const int64_t n4 = n_ref >> 2;
        for (int64_t b0 = 0; b0 < n4; b0 += 64 * kCrossBatch) {
            float4 d[kCrossBatch], a[kCrossBatch];
#pragma unroll
            for (int u = 0; u < kCrossBatch; ++u) {
                const int64_t j = b0 + u * 64 + lane;
                d[u] = j < n4 ? row4[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                a[u] = j < n4 ? an4[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
#pragma unroll
            for (int u = 0; u < kCrossBatch; ++u) {
                const int64_t j = b0 + u * 64 + lane;
                if (j < n4) {
                    const uint32_t r = (uint32_t)(4 * j) + ref_offset;
                    topk_insert(lst, pack_key(fmaf(g2, a[u].x, fmaf(m2g, d[u].x, bni)), r));
                    topk_insert(lst, pack_key(fmaf(g2, a[u].y, fmaf(m2g, d[u].y, bni)), r + 1));
                    topk_insert(lst, pack_key(fmaf(g2, a[u].z, fmaf(m2g, d[u].z, bni)), r + 2));
                    topk_insert(lst, pack_key(fmaf(g2, a[u].w, fmaf(m2g, d[u].w, bni)), r + 3));
                }
            }
        }
    } else {
        constexpr int U = 4 * kCrossBatch;
        for (int64_t b0 = 0; b0 < n_ref; b0 += 64 * U) {
            float d[U], a[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t r = b0 + u * 64 + lane;
                d[u] = r < n_ref ? row[r] : 0.0f;
                a[u] = r < n_ref ? an[r] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t r = b0 + u * 64 + lane;
                if (r < n_ref) topk_insert(lst, pack_key(fmaf(g2, a[u], fmaf(m2g, d[u], bni)), (uint32_t)r + ref_offset));
            }
        }
    }
    // k rounds: the wave minimum of the list heads; its lane pops it
    for (int s = 0; s < k; ++s) {
        unsigned long long m = lst[0];
        for (int off = 32; off > 0; off >>= 1) m = umin64(m, shfl_xor_u64(m, off));
        if (lst[0] == m && m != kEmpty) {
#pragma unroll
            for (int q = 0; q < L - 1; ++q) lst[q] = lst[q + 1];
            lst[L - 1] = kEmpty;
        }
        if (lane == 0) slots[s] = m;
    }
}

// ---------------------------------------------------------------- hipBLASLt cross term
struct BlasLt {
    hipblasLtHandle_t handle = nullptr;
    void* workspace = nullptr;
    size_t ws_size = 0;
    std::map<std::tuple<int64_t, int64_t, int64_t>, hipblasLtMatmulAlgo_t> algos;
};
std::mutex g_blas_mutex;
std::map<int, BlasLt> g_blas;  // per device; kept until the process ends

constexpr size_t kBlasWorkspace = (size_t)64 << 20;
constexpr int kBlasCandidates = 8;

int blas_fail(const char* what, hipblasStatus_t st) {
    set_error("match_cross_bf16: %s failed (hipblasStatus %d)", what, (int)st);
    return CMTIP_ERR_LAUNCH;
}

}  // namespace

extern "C" int cmtip_match_cross_bf16(const uint16_t* A, int64_t n_ref, const uint16_t* B, int64_t n_img, int64_t K,
                                      float* Dt, void* stream) {
    if (n_ref < 0 || n_img < 0 || K < 1 || (n_ref > 0 && n_img > 0 && (!A || !B || !Dt))) {
        set_error("match_cross_bf16: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (n_ref == 0 || n_img == 0) return CMTIP_OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("match_cross_bf16: hipGetDevice failed");
        return CMTIP_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(g_blas_mutex);
    BlasLt& bl = g_blas[dev];
    hipblasStatus_t st;
    if (!bl.handle) {
        if ((st = hipblasLtCreate(&bl.handle)) != HIPBLAS_STATUS_SUCCESS) return blas_fail("hipblasLtCreate", st);
        if (hipMalloc(&bl.workspace, kBlasWorkspace) != hipSuccess) {
            set_error("match_cross_bf16: workspace allocation failed");
            return CMTIP_ERR_ALLOC;
        }
        bl.ws_size = kBlasWorkspace;
    }
    // column-major view: Dt (n_ref x n_img, ld n_ref) = A^T (n_ref x K) * B (K x n_img),
    // i.e. row-major Dt[i][r] = <A_r, B_i> (A, B row-major with K columns)
    hipblasLtMatmulDesc_t desc = nullptr;
    hipblasLtMatrixLayout_t la = nullptr, lb = nullptr, lc = nullptr;
    int rc = CMTIP_OK;
    const hipblasOperation_t opA = HIPBLAS_OP_T, opB = HIPBLAS_OP_N;
    do {
        if ((st = hipblasLtMatmulDescCreate(&desc, HIPBLAS_COMPUTE_32F, HIP_R_32F)) != HIPBLAS_STATUS_SUCCESS ||
            (st = hipblasLtMatmulDescSetAttribute(desc, HIPBLASLT_MATMUL_DESC_TRANSA, &opA, sizeof(opA))) !=
                HIPBLAS_STATUS_SUCCESS ||
            (st = hipblasLtMatmulDescSetAttribute(desc, HIPBLASLT_MATMUL_DESC_TRANSB, &opB, sizeof(opB))) !=
                HIPBLAS_STATUS_SUCCESS) {
            rc = blas_fail("matmul descriptor", st);
            break;
        }
        if ((st = hipblasLtMatrixLayoutCreate(&la, HIP_R_16BF, K, n_ref, K)) != HIPBLAS_STATUS_SUCCESS ||
            (st = hipblasLtMatrixLayoutCreate(&lb, HIP_R_16BF, K, n_img, K)) != HIPBLAS_STATUS_SUCCESS ||
            (st = hipblasLtMatrixLayoutCreate(&lc, HIP_R_32F, n_ref, n_img, n_ref)) != HIPBLAS_STATUS_SUCCESS) {
            rc = blas_fail("matrix layouts", st);
            break;
        }
        const float alpha = 1.0f, beta = 0.0f;
        const auto key = std::make_tuple(n_ref, n_img, K);
        auto it = bl.algos.find(key);
        if (it == bl.algos.end()) {
            // first call for this shape: time up to kBlasCandidates heuristic algorithms
            // once each (HIP events on the stream; one synchronisation) and keep the fastest
            hipblasLtMatmulPreference_t pref = nullptr;
            hipblasLtMatmulHeuristicResult_t heur[kBlasCandidates];
            int found = 0;
            const uint64_t ws = bl.ws_size;
            st = hipblasLtMatmulPreferenceCreate(&pref);
            if (st == HIPBLAS_STATUS_SUCCESS)
                st = hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &ws,
                                                           sizeof(ws));
            if (st == HIPBLAS_STATUS_SUCCESS)
                st = hipblasLtMatmulAlgoGetHeuristic(bl.handle, desc, la, lb, lc, lc, pref, kBlasCandidates, heur,
                                                     &found);
            if (pref) (void)hipblasLtMatmulPreferenceDestroy(pref);
            if (st != HIPBLAS_STATUS_SUCCESS || found < 1) {
                rc = blas_fail("algorithm heuristic", st);
                break;
            }
            int pick = 0;
            if (found > 1) {
                hipEvent_t e0 = nullptr, e1 = nullptr;
                float best_ms = 0.f;
                bool have = false;
                // per candidate: an untimed launch (its first launch loads its code), then a
                // timed one
                if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
                    for (int a = 0; a < found; ++a) {
                        if (heur[a].state != HIPBLAS_STATUS_SUCCESS || heur[a].workspaceSize > bl.ws_size) continue;
                        float ms = 0.f;
                        if (hipblasLtMatmul(bl.handle, desc, &alpha, A, la, B, lb, &beta, Dt, lc, Dt, lc,
                                            &heur[a].algo, bl.workspace, bl.ws_size,
                                            as_stream(stream)) != HIPBLAS_STATUS_SUCCESS ||
                            hipEventRecord(e0, as_stream(stream)) != hipSuccess ||
                            hipblasLtMatmul(bl.handle, desc, &alpha, A, la, B, lb, &beta, Dt, lc, Dt, lc,
                                            &heur[a].algo, bl.workspace, bl.ws_size,
                                            as_stream(stream)) != HIPBLAS_STATUS_SUCCESS ||
                            hipEventRecord(e1, as_stream(stream)) != hipSuccess ||
                            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                            continue;
                        if (!have || ms < best_ms) {
                            have = true;
                            best_ms = ms;
                            pick = a;
                        }
                    }
                }
                if (e0) (void)hipEventDestroy(e0);
                if (e1) (void)hipEventDestroy(e1);
            }
            it = bl.algos.emplace(key, heur[pick].algo).first;
        }
        st = hipblasLtMatmul(bl.handle, desc, &alpha, A, la, B, lb, &beta, Dt, lc, Dt, lc, &it->second, bl.workspace,
                             bl.ws_size, as_stream(stream));
        if (st != HIPBLAS_STATUS_SUCCESS) rc = blas_fail("hipblasLtMatmul", st);
    } while (false);
    if (la) (void)hipblasLtMatrixLayoutDestroy(la);
    if (lb) (void)hipblasLtMatrixLayoutDestroy(lb);
    if (lc) (void)hipblasLtMatrixLayoutDestroy(lc);
    if (desc) (void)hipblasLtMatmulDescDestroy(desc);
    return rc;
}

extern "C" int cmtip_match_topk_cross(const float* Dt, int64_t n_ref, int64_t ref_offset, int64_t n_img,
                                      const float* an, const float* bn, const double* gamma, int k,
                                      unsigned long long* cand, void* stream) {
    if (k < 1 || k > kTopKMax) {
        set_error("match_topk_cross: k must be in [1, %d]", kTopKMax);
        return CMTIP_ERR_ARG;
    }
    if (n_ref < 0 || n_img < 0 || ref_offset < 0 || ref_offset + n_ref > 0xffffffffLL ||
        (n_ref > 0 && n_img > 0 && (!Dt || !an || !bn || !cand))) {
        set_error("match_topk_cross: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (n_ref == 0 || n_img == 0) return CMTIP_OK;
    const int64_t blocks = (n_img + kCrossWaves - 1) / kCrossWaves;
    if (blocks > 0x7fffffffLL) {
        set_error("match_topk_cross: too many images");
        return CMTIP_ERR_ARG;
    }
    const bool vec = n_ref % 4 == 0 && aligned16(Dt) && aligned16(an);
#define CMTIP_TOPK_GO(V, L)                                                                                    \
    hipLaunchKernelGGL((topk_cross_kernel<V, L>), dim3((unsigned)blocks), dim3(64 * kCrossWaves), 0,           \
                       as_stream(stream), Dt, n_ref, (uint32_t)ref_offset, n_img, an, bn, gamma, k, cand)
    // each lane keeps a list as long as k, so the merged top k is exact
    if (k <= kTopK) {
        if (vec) CMTIP_TOPK_GO(true, kTopK); else CMTIP_TOPK_GO(false, kTopK);
    } else {
        if (vec) CMTIP_TOPK_GO(true, kTopKMax); else CMTIP_TOPK_GO(false, kTopKMax);
    }
#undef CMTIP_TOPK_GO
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

namespace {

// Launchers: one workgroup per tile (the LDS of one tile fills a CU), dynamic LDS raised
// to what the kernel needs; L = 8 list entries for k <= 8, else 16.
template <int WM, int WN, int TM, int TN>
struct XTileLaunch {
    using T = XTile<WM, WN, TM, TN>;
    static int64_t tiles_m(int64_t n_ref) { return (n_ref + T::BMt - 1) / T::BMt; }
    static int64_t tiles(int64_t n_ref, int64_t n_img) { return tiles_m(n_ref) * ((n_img + T::BNt - 1) / T::BNt); }
    template <int L, bool AL>
    static void go(const uint16_t* A, int64_t n_ref, uint32_t off, const uint16_t* Bw, int64_t n_img, int64_t P,
                   const float* an, const float* bn, const double* gamma, int k, unsigned long long* cand,
                   hipStream_t st) {
        static bool attr = false;
        if (!attr) {
            (void)hipFuncSetAttribute((const void*)cross_topk_bf16_kernel<WM, WN, TM, TN, L, AL>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)T::kLds);
            attr = true;
        }
        hipLaunchKernelGGL((cross_topk_bf16_kernel<WM, WN, TM, TN, L, AL>), dim3((unsigned)tiles(n_ref, n_img)),
                           dim3(T::kThreads), T::kLds, st, A, n_ref, off, Bw, n_img, P, an, bn, gamma, k, cand,
                           (int)tiles_m(n_ref));
    }
    static void run(bool aligned, const uint16_t* A, int64_t n_ref, uint32_t off, const uint16_t* Bw, int64_t n_img,
                    int64_t P, const float* an, const float* bn, const double* gamma, int k,
                    unsigned long long* cand, hipStream_t st) {
        if (k <= kTopK) {
            if (aligned) go<kTopK, true>(A, n_ref, off, Bw, n_img, P, an, bn, gamma, k, cand, st);
            else go<kTopK, false>(A, n_ref, off, Bw, n_img, P, an, bn, gamma, k, cand, st);
        } else {
            if (aligned) go<kTopKMax, true>(A, n_ref, off, Bw, n_img, P, an, bn, gamma, k, cand, st);
            else go<kTopKMax, false>(A, n_ref, off, Bw, n_img, P, an, bn, gamma, k, cand, st);
        }
    }
};

template <int WM, int WN, int TM, int TN, int S>
struct DTileLaunch {
    using T = DTile<WM, WN, TM, TN, S>;
    static int64_t tiles_m(int64_t n_ref) { return (n_ref + T::BMt - 1) / T::BMt; }
    static int64_t tiles(int64_t n_ref, int64_t n_img) { return tiles_m(n_ref) * ((n_img + T::BNt - 1) / T::BNt); }
    template <int L>
    static void go(const uint16_t* A, int64_t n_ref, uint32_t off, const uint16_t* Bw, int64_t n_img, int64_t P,
                   const float* an, const float* bn, const double* gamma, int k, unsigned long long* cand,
                   hipStream_t st) {
        static bool attr = false;
        if (!attr) {
            (void)hipFuncSetAttribute((const void*)cross_topk_bf16_dma_kernel<WM, WN, TM, TN, S, L>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)T::kLds);
            attr = true;
        }
        hipLaunchKernelGGL((cross_topk_bf16_dma_kernel<WM, WN, TM, TN, S, L>), dim3((unsigned)tiles(n_ref, n_img)),
                           dim3(T::kThreads), T::kLds, st, A, n_ref, off, Bw, n_img, P, an, bn, gamma, k, cand,
                           (int)tiles_m(n_ref));
    }
    static void run(const uint16_t* A, int64_t n_ref, uint32_t off, const uint16_t* Bw, int64_t n_img, int64_t P,
                    const float* an, const float* bn, const double* gamma, int k, unsigned long long* cand,
                    hipStream_t st) {
        if (k <= kTopK) go<kTopK>(A, n_ref, off, Bw, n_img, P, an, bn, gamma, k, cand, st);
        else go<kTopKMax>(A, n_ref, off, Bw, n_img, P, an, bn, gamma, k, cand, st);
    }
};

// 256 x 256 tiles for both: measured at 10000 x 5000 x 16384, the register-staged kernel
// 2.25-2.31 ms with 256 x 256, 2.15-2.16 with 256 x 160 or 256 x 192 tiles (8 waves);
// the direct-to-LDS kernel 1.80 ms (four waves of 128 x 128; eight waves of 128 x 64,
// 1.76-1.99 ms depending on the stage count, spill at five stages).
using XTileBF16 = XTileLaunch<2, 4, 8, 4>;
using DTileBF16 = DTileLaunch<2, 2, 8, 8, 4>;

}  // namespace

extern "C" int cmtip_match_topk_bf16(const uint16_t* A, int64_t n_ref, int64_t ref_offset, const uint16_t* Bw,
                                     int64_t n_img, int64_t P, const float* an, const float* bn,
                                     const double* gamma, int k, unsigned long long* cand, void* stream) {
    if (k < 1 || k > kTopKMax) {
        set_error("match_topk_bf16: k must be in [1, %d]", kTopKMax);
        return CMTIP_ERR_ARG;
    }
    int64_t tiles_m, tiles;
    CMTIP_TRY(check_match_args(n_ref, ref_offset, n_img, P, A, Bw, an, bn, cand, &tiles_m, &tiles));
    if (tiles == 0) return CMTIP_OK;
    const bool aligned = P % 8 == 0 && aligned16(A) && aligned16(Bw);
    const hipStream_t st = as_stream(stream);
    if (aligned && P % (2 * BK32) == 0)
        DTileBF16::run(A, n_ref, (uint32_t)ref_offset, Bw, n_img, P, an, bn, gamma, k, cand, st);
    else
        XTileBF16::run(aligned, A, n_ref, (uint32_t)ref_offset, Bw, n_img, P, an, bn, gamma, k, cand, st);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_match_rescore(const float* A_all, int64_t n_ref_total, const float* B, int64_t n_img,
                                   int64_t P, const float* w2, const double* gamma_of_ref, int k,
                                   const unsigned long long* cand, unsigned long long* best, void* stream) {
    return cmtip_match_rescore_masked(A_all, n_ref_total, B, nullptr, n_img, P, w2, gamma_of_ref, k, cand, best,
                                      stream);
}

extern "C" int cmtip_match_rescore_masked(const float* A_all, int64_t n_ref_total, const float* B,
                                          const uint8_t* mask, int64_t n_img, int64_t P, const float* w2,
                                          const double* gamma_of_ref, int k, const unsigned long long* cand,
                                          unsigned long long* best, void* stream) {
    if (k < 1 || k > kTopKMax || n_img < 0 || P < 1 || n_ref_total < 0 ||
        (n_img > 0 && (!A_all || !B || !w2 || !gamma_of_ref || !cand || !best))) {
        set_error("match_rescore: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (n_img == 0) return CMTIP_OK;
    if (n_img > 0x7fffffff) {
        set_error("match_rescore: too many images");
        return CMTIP_ERR_ARG;
    }
    const bool quads = P % 4 == 0 && ((reinterpret_cast<uintptr_t>(A_all) | reinterpret_cast<uintptr_t>(B) |
                                        reinterpret_cast<uintptr_t>(w2)) & 15) == 0 &&
                       (reinterpret_cast<uintptr_t>(mask) & 3) == 0;
    if (quads && k <= kTopK)
        hipLaunchKernelGGL(rescore4_kernel<kTopK>, dim3((unsigned)n_img), dim3(kThreads), 0, as_stream(stream), A_all,
                           n_ref_total, B, P, w2, gamma_of_ref, k, cand, best, mask);
    else if (quads)
        hipLaunchKernelGGL(rescore4_kernel<kTopKMax>, dim3((unsigned)n_img), dim3(kThreads), 0, as_stream(stream),
                           A_all, n_ref_total, B, P, w2, gamma_of_ref, k, cand, best, mask);
    else
        hipLaunchKernelGGL(rescore_kernel, dim3((unsigned)n_img), dim3(kThreads), 0, as_stream(stream), A_all,
                           n_ref_total, B, P, w2, gamma_of_ref, k, cand, best, mask);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_image_stats_masked(const float* X, const uint8_t* mask, int64_t rows, int64_t P, double* colsum,
                                        float* c, double* mom, void* stream) {
    if (rows < 0 || P < 1 || !colsum || (rows > 0 && (!X || !mask))) {
        set_error("image_stats_masked: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(column_sums_masked_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64 * kStatWaves), 0, st, X,
                       mask, rows, P, colsum);
    hipLaunchKernelGGL(finish_stats_masked_kernel, dim3(1), dim3(kThreads), 0, st, (const double*)colsum, P, c, mom);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_match_prepare_masked(const float* B, const uint8_t* mask, int64_t n_img, int64_t P,
                                          const float* w2, const float* c, float* Bw2, uint16_t* Bw2_bf16, float* bn,
                                          void* stream) {
    if (n_img < 0 || P < 1 || n_img > 0x7fffffff ||
        (n_img > 0 && (!B || !mask || !w2 || !c || !bn || (!Bw2 && !Bw2_bf16)))) {
        set_error("match_prepare_masked: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (n_img == 0) return CMTIP_OK;
    hipLaunchKernelGGL(prepare_masked_kernel, dim3((unsigned)n_img), dim3(kThreads), 0, as_stream(stream), B, mask, P,
                       w2, c, Bw2, Bw2_bf16, bn);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_center_rows_sq(const float* X, int64_t rows, int64_t P, const double* gamma, const float* c,
                                    float* Y2, uint16_t* Y2_bf16, void* stream) {
    if (rows < 0 || P < 1 || rows > 0x7fffffff || (rows > 0 && (!X || !c || (!Y2 && !Y2_bf16)))) {
        set_error("center_rows_sq: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (rows == 0) return CMTIP_OK;
    hipLaunchKernelGGL(center_rows_sq_kernel, dim3((unsigned)rows), dim3(kThreads), 0, as_stream(stream), X, P, gamma,
                       c, Y2, Y2_bf16);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}
