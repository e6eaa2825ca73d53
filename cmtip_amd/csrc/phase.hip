// Phase retrieval of an oversampled autocorrelation on the device: cmtip/phasing/phase.py.
//
// All volumes are M^3 float32 in FFT position order (the reference's trailing-underscore
// arrays, DC at index 0) unless a comment says "centred".  One ER or HIO step
// (phase.py:83-144) is five FFT axis passes over a complex M^3 scratch volume:
//   axis 2 forward (real density in), axis 1 forward,
//   axis 0 forward -> amplitude replacement -> axis 0 inverse (one LDS-resident pass),
//   axis 1 inverse, axis 2 inverse with 1/M^3, the real part and the ER / HIO update of
//   the density folded into the store.
// The amplitudes carry the DC mask of phase.py:194-195 as a negative entry (the value
// is kept there).  Gaussian filters (scipy.ndimage.gaussian_filter, modes 'constant'
// and 'wrap') are separable FP64 passes in scipy's order of operations; maxima use
// integer atomics on the bits of non-negative values, so every result here is
// independent of scheduling.
#include <math.h>

#include <initializer_list>
#include <vector>

#include "fft.h"

using namespace cmtip;

struct cmtip_phaser_s {
    int M = 0;
    Fft1D fft;
    DevBuf amp;    // M^3 float amplitudes (negative = keep the value)
    DevBuf t;      // M^3 float2 scratch volume
    DevBuf da, db; // M^3 double filter buffers
    DevBuf red;    // reduction scratch: kRed x 4 doubles + 1 u64 max
};

namespace {

constexpr int kT = 256;
constexpr int kRed = 512;      // blocks of the two-stage sums
constexpr int kMaxRadius = 8;  // gaussian_filter radius int(truncate * sigma + 0.5)

unsigned nblk(int64_t n, int64_t cap = 8192) {
    int64_t b = (n + kT - 1) / kT;
    if (b > cap) b = cap;
    return (unsigned)(b < 1 ? 1 : b);
}

bool valid(const cmtip_phaser_s* p) {
    if (!p || p->M < 1) {
        set_error("phaser: null or destroyed object");
        return false;
    }
    return true;
}

// ---------------------------------------------------------------- gaussian filter
struct GaussW {
    double w[kMaxRadius + 1];
    int r;
};

// scipy's _gaussian_kernel1d: x = arange(-r, r + 1), phi = exp(-0.5 / sigma^2 * x^2),
// phi / phi.sum(); the filter is symmetric, so only w[0..r] are kept.
int gauss_weights(double sigma, double truncate, GaussW* g) {
    const int r = (int)(truncate * sigma + 0.5);
    if (!(sigma > 0.0) || r < 0 || r > kMaxRadius) {
        set_error("phase: gaussian filter radius %d out of range (sigma %g)", r, sigma);
        return CMTIP_ERR_ARG;
    }
    std::vector<double> phi(2 * r + 1);
    const double s2 = sigma * sigma;
    for (int i = 0; i <= 2 * r; ++i) {
        const double x = (double)(i - r);
        phi[i] = exp(-0.5 / s2 * (x * x));
    }
    double sum = 0.0;
    for (double v : phi) sum += v;
    for (int j = 0; j <= r; ++j) g->w[j] = phi[r + j] / sum;
    g->r = r;
    return CMTIP_OK;
}

__device__ __forceinline__ int wrap_i(int i, int M) {
    i %= M;
    return i < 0 ? i + M : i;
}

// One axis of scipy's correlate1d with a symmetric filter: w0 x[i] + sum_j wj (x[i-j] +
// x[i+j]), the farthest pair first as scipy adds them; mode 0 'constant' (zero outside), 1 'wrap' (periodic).  The first axis can
// read float input (src_f, with an optional max(x, 0) or |x| map) instead of double.
__global__ void gauss_axis_kernel(const float* __restrict__ src_f, const double* __restrict__ src_d,
                                  double* __restrict__ out, int M, int axis, GaussW g, int mode, int map) {
#pragma clang fp contract(off)
    const int64_t M3 = (int64_t)M * M * M;
    const int64_t stride = axis == 0 ? (int64_t)M * M : (axis == 1 ? M : 1);
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < M3; t += (int64_t)gridDim.x * kT) {
        const int i = (int)((t / stride) % M);
        const int64_t base = t - (int64_t)i * stride;
        auto at = [&](int j) -> double {
            if (mode == 0) {
                if (j < 0 || j >= M) return 0.0;
            } else {
                j = wrap_i(j, M);
            }
            const int64_t o = base + (int64_t)j * stride;
            if (src_f) {
                const float v = src_f[o];
                return map == 1 ? (double)fmaxf(v, 0.f) : (map == 2 ? (double)fabsf(v) : (double)v);
            }
            return src_d[o];
        };
        double v = at(i) * g.w[0];
        for (int j = g.r; j >= 1; --j) v += (at(i - j) + at(i + j)) * g.w[j];
        out[t] = v;
    }
}

// max of non-negative doubles (their bit patterns order like the values)
__global__ void max_nonneg_kernel(const double* __restrict__ src_d, const float* __restrict__ src_f, int absf,
                                  int64_t n, unsigned long long* __restrict__ out) {
    double m = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < n; t += (int64_t)gridDim.x * kT) {
        const double v = src_d ? src_d[t] : (absf ? (double)fabsf(src_f[t]) : (double)src_f[t]);
        m = v > m ? v : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

__global__ void fill_u64_kernel(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// float copy of a centred double volume, ifftshifted to FFT position order
__global__ void shift_to_float_kernel(const double* __restrict__ src, float* __restrict__ dst, int M) {
    const int64_t M3 = (int64_t)M * M * M;
    const int h = M / 2;
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < M3; t += (int64_t)gridDim.x * kT) {
        const int i0 = (int)(t / ((int64_t)M * M)), i1 = (int)((t / M) % M), i2 = (int)(t % M);
        const int64_t s = ((int64_t)((i0 + h) % M) * M + (i1 + h) % M) * M + (i2 + h) % M;
        dst[t] = (float)src[s];
    }
}

// amplitudes_ = sqrt(|F|) with the DC entry marked kept (phase.py:190-195)
__global__ void sqrt_amp_kernel(float* __restrict__ a, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < n; t += (int64_t)gridDim.x * kT)
        a[t] = t == 0 ? -1.0f : sqrtf(a[t]);
}

// create_support_ (phase.py:47-64): the centred cube [Mquat, M - Mquat)^3 ifftshifted
// (empty for Mquat = 0: the slice(0, -0) of :57 selects nothing), and
// ac_filt_ > 1e-2 max(ac_filt_), from the centred FP64 filtered volume
__global__ void create_support_kernel(const double* __restrict__ acd, const unsigned long long* __restrict__ mx,
                                      float* __restrict__ support, int M, int Mquat) {
#pragma clang fp contract(off)
    const int64_t M3 = (int64_t)M * M * M;
    const int h = M / 2;
    const double thr = 1e-2 * __longlong_as_double((long long)*mx);
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < M3; t += (int64_t)gridDim.x * kT) {
        const int c0 = ((int)(t / ((int64_t)M * M)) + h) % M, c1 = ((int)((t / M) % M) + h) % M;
        const int c2 = ((int)(t % M) + h) % M;
        const bool cube = Mquat > 0 && c0 >= Mquat && c0 < M - Mquat && c1 >= Mquat && c1 < M - Mquat && c2 >= Mquat &&
                          c2 < M - Mquat;
        const double v = acd[((int64_t)c0 * M + c1) * M + c2];
        support[t] = (cube && v > thr) ? 1.0f : 0.0f;
    }
}

// shrink-wrap threshold (phase.py:162): support = gauss(|rho|) > max|rho| * cutoff
__global__ void threshold_kernel(const double* __restrict__ g, const unsigned long long* __restrict__ mx,
                                 double cutoff, float* __restrict__ support, int64_t n) {
#pragma clang fp contract(off)
    const double thr = __longlong_as_double((long long)*mx) * cutoff;
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < n; t += (int64_t)gridDim.x * kT)
        support[t] = g[t] > thr ? 1.0f : 0.0f;
}

// center_of_mass sums (phase.py:11-23): per-block partials of |rho|, |rho| h_d with the
// ifftshifted cell-centre coordinates ls[(i + M/2) % M], ls[j] = -1 + (2j + 1) / M
__global__ void com_partial_kernel(const float* __restrict__ rho, int M, double* __restrict__ part) {
    __shared__ double red[4][kT];
    const int64_t M3 = (int64_t)M * M * M;
    const int h = M / 2;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < M3; t += (int64_t)gridDim.x * kT) {
        const double a = fabs((double)rho[t]);
        const int i[3] = {(int)(t / ((int64_t)M * M)), (int)((t / M) % M), (int)(t % M)};
        s[0] += a;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int j = (i[d] + h) % M;
            const double c = (-1.0 + 1.0 / M) + 2.0 * (double)j / (double)M;
            s[d + 1] += a * c;
        }
    }
#pragma unroll
    for (int d = 0; d < 4; ++d) red[d][threadIdx.x] = s[d];
    __syncthreads();
    for (int k = kT / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k)
#pragma unroll
            for (int d = 0; d < 4; ++d) red[d][threadIdx.x] += red[d][threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void com_finish_kernel(const double* __restrict__ part, int nb, double* __restrict__ out) {
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += part[b * 4 + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// np.roll by -s along every axis: out[i] = in[(i + s) mod M]
__global__ void roll_kernel(const float* __restrict__ in, float* __restrict__ out, int M, int s0, int s1, int s2) {
    const int64_t M3 = (int64_t)M * M * M;
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < M3; t += (int64_t)gridDim.x * kT) {
        const int i0 = (int)(t / ((int64_t)M * M)), i1 = (int)((t / M) % M), i2 = (int)(t % M);
        const int64_t src = ((int64_t)wrap_i(i0 + s0, M) * M + wrap_i(i1 + s1, M)) * M + wrap_i(i2 + s2, M);
        out[t] = in[src];
    }
}

__global__ void copy_kernel(const float* __restrict__ in, float* __restrict__ out, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < n; t += (int64_t)gridDim.x * kT) out[t] = in[t];
}

// ---------------------------------------------------------------- FFT pass sequences
// Line enumerations of the three axes of an M^3 volume (u, v: the other two axes, v
// contiguous when the transform axis is strided).
void axis_lines(AxisDesc& d, int M, int axis, bool in_side, bool out_side) {
    const int64_t M2 = (int64_t)M * M;
    int64_t su, sv, sa;
    if (axis == 2) {
        su = M2; sv = M; sa = 1;
    } else if (axis == 1) {
        su = M2; sv = 1; sa = M;
    } else {
        su = M; sv = 1; sa = M2;
    }
    if (in_side) {
        d.in_su = su; d.in_sv = sv; d.in_sa = sa;
        d.ia = M; d.ig1 = 0;
    }
    if (out_side) {
        d.out_su = su; d.out_sv = sv; d.out_sa = sa;
        d.oa = M; d.og1 = 0;
    }
    d.nu = M;
    d.nv = M;
}

// forward passes on axes 2 and 1 of a real volume into the complex scratch
int forward_21(cmtip_phaser_s* p, const float* rho, hipStream_t st) {
    const int M = p->M;
    float2* t = p->t.as<float2>();
    AxisDesc d;
    axis_lines(d, M, 2, true, true);
    d.in = rho;
    d.in_real = 1;
    d.out = t;
    d.sign = -1;
    CMTIP_TRY(fft_axis_pass(p->fft, d, st));
    d = AxisDesc();
    axis_lines(d, M, 1, true, true);
    d.in = t;
    d.out = t;
    d.sign = -1;
    return fft_axis_pass(p->fft, d, st);
}

// output positions p -> fftshift order along the pass axis
void shift_out(AxisDesc& d, int M) {
    d.oa = M - M / 2;
    d.og1 = M / 2;
    d.ob = M / 2;
    d.og2 = 0;
}

// one ER (mode 1) or HIO (mode 2) step on rho in place; mode 0 writes the projected
// density rho_mod itself
int phase_step(cmtip_phaser_s* p, float* rho, const float* support, int mode, float beta, float rho_max,
               hipStream_t st) {
    const int M = p->M;
    float2* t = p->t.as<float2>();
    CMTIP_TRY(forward_21(p, rho, st));
    AxisDesc d;
    axis_lines(d, M, 0, true, true);
    d.in = t;
    d.out = t;
    d.sign = -1;
    d.mid_amp = p->amp.as<const float>();
    d.mid_su = M;
    d.mid_sv = 1;
    d.mid_sa = (int64_t)M * M;
    CMTIP_TRY(fft_axis_pass(p->fft, d, st));
    d = AxisDesc();
    axis_lines(d, M, 1, true, true);
    d.in = t;
    d.out = t;
    d.sign = 1;
    CMTIP_TRY(fft_axis_pass(p->fft, d, st));
    d = AxisDesc();
    axis_lines(d, M, 2, true, true);
    d.in = t;
    d.out = rho;
    d.out_mode = 1;
    d.out_scale = (float)(1.0 / ((double)M * M * M));
    d.sign = 1;
    d.upd_mode = mode;
    d.upd_support = support;
    d.upd_beta = beta;
    d.upd_rho_max = rho_max;
    return fft_axis_pass(p->fft, d, st);
}

int gauss3(cmtip_phaser_s* p, const float* src, int map, double sigma, double truncate, int mode, hipStream_t st) {
    GaussW g;
    CMTIP_TRY(gauss_weights(sigma, truncate, &g));
    const int M = p->M;
    const int64_t M3 = (int64_t)M * M * M;
    double* a = p->da.as<double>();
    double* b = p->db.as<double>();
    const unsigned gr = nblk(M3);
    hipLaunchKernelGGL(gauss_axis_kernel, dim3(gr), dim3(kT), 0, st, src, (const double*)nullptr, a, M, 0, g, mode,
                       map);
    hipLaunchKernelGGL(gauss_axis_kernel, dim3(gr), dim3(kT), 0, st, (const float*)nullptr, (const double*)a, b, M, 1,
                       g, mode, 0);
    hipLaunchKernelGGL(gauss_axis_kernel, dim3(gr), dim3(kT), 0, st, (const float*)nullptr, (const double*)b, a, M, 2,
                       g, mode, 0);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;  // result in da
}

unsigned long long* max_slot(cmtip_phaser_s* p) { return p->red.as<unsigned long long>() + 4 * kRed + 4; }

int max_into_slot(cmtip_phaser_s* p, const double* src_d, const float* src_f, int absf, hipStream_t st) {
    const int64_t M3 = (int64_t)p->M * p->M * p->M;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, st, max_slot(p), 1);
    hipLaunchKernelGGL(max_nonneg_kernel, dim3(nblk(M3, 1024)), dim3(kT), 0, st, src_d, src_f, absf, M3, max_slot(p));
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

}  // namespace

// ---------------------------------------------------------------- C ABI

extern "C" int cmtip_phaser_create(int M, cmtip_phaser* out) {
    if (!out) {
        set_error("phaser_create: null output");
        return CMTIP_ERR_ARG;
    }
    *out = nullptr;
    if (M < 2 || M > 1024) {
        set_error("phaser_create: M = %d out of range [2, 1024]", M);
        return CMTIP_ERR_ARG;
    }
    cmtip_phaser_s* p = new cmtip_phaser_s();
    p->M = M;
    const int64_t M3 = (int64_t)M * M * M;
    int rc = fft1d_init(p->fft, M);
    if (rc == CMTIP_OK) rc = p->amp.ensure(sizeof(float) * (size_t)M3);
    if (rc == CMTIP_OK) rc = p->t.ensure(sizeof(float2) * (size_t)M3);
    if (rc == CMTIP_OK) rc = p->da.ensure(sizeof(double) * (size_t)M3);
    if (rc == CMTIP_OK) rc = p->db.ensure(sizeof(double) * (size_t)M3);
    if (rc == CMTIP_OK) rc = p->red.ensure(sizeof(double) * (4 * kRed + 8));
    if (rc != CMTIP_OK) {
        cmtip_phaser_destroy(p);
        return rc;
    }
    *out = p;
    return CMTIP_OK;
}

extern "C" int cmtip_phaser_destroy(cmtip_phaser p) {
    if (!p) return CMTIP_OK;
    fft1d_free(p->fft);
    p->amp.release();
    p->t.release();
    p->da.release();
    p->db.release();
    p->red.release();
    p->M = 0;
    delete p;
    return CMTIP_OK;
}

extern "C" int cmtip_phase_setup(cmtip_phaser p, const float* ac, float* ac_filt_, float* support_, int Mquat,
                                 void* stream) {
    if (!valid(p)) return CMTIP_ERR_ARG;
    const int M = p->M;
    if (!ac || (support_ && (Mquat < 0 || 2 * Mquat > M))) {
        set_error("phase_setup: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipStream_t st = as_stream(stream);
    const int64_t M3 = (int64_t)M * M * M;
    // ac_filt = gaussian_filter(max(ac, 0), 'constant', sigma 1, truncate 2) (phase.py:186-187)
    CMTIP_TRY(gauss3(p, ac, 1, 1.0, 2.0, 0, st));
    const double* acd = p->da.as<const double>();
    // float ac_filt_ (ifftshifted): the caller's buffer, else the second filter buffer
    // (free once the filter is done; M^3 floats fit in its M^3 doubles)
    float* f = ac_filt_ ? ac_filt_ : reinterpret_cast<float*>(p->db.p);
    hipLaunchKernelGGL(shift_to_float_kernel, dim3(nblk(M3)), dim3(kT), 0, st, acd, f, M);
    CMTIP_LAUNCH_CHECK();
    if (support_) {
        CMTIP_TRY(max_into_slot(p, acd, nullptr, 0, st));
        hipLaunchKernelGGL(create_support_kernel, dim3(nblk(M3)), dim3(kT), 0, st, acd,
                           (const unsigned long long*)max_slot(p), support_, M, Mquat);
        CMTIP_LAUNCH_CHECK();
    }
    // amplitudes_ = sqrt(|fftn(ac_filt_)|) (phase.py:190-192): axes 2, 1 forward into the
    // scratch volume, axis 0 forward writing |F| into the amplitude array
    CMTIP_TRY(forward_21(p, f, st));
    AxisDesc d;
    axis_lines(d, M, 0, true, true);
    d.in = p->t.p;
    d.out = p->amp.p;
    d.out_mode = 3;
    d.sign = -1;
    CMTIP_TRY(fft_axis_pass(p->fft, d, st));
    hipLaunchKernelGGL(sqrt_amp_kernel, dim3(nblk(M3)), dim3(kT), 0, st, p->amp.as<float>(), M3);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_phase_get_amplitudes(cmtip_phaser p, float* amplitudes_, void* stream) {
    if (!valid(p)) return CMTIP_ERR_ARG;
    if (!amplitudes_) {
        set_error("phase_get_amplitudes: null output");
        return CMTIP_ERR_ARG;
    }
    const int64_t M3 = (int64_t)p->M * p->M * p->M;
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(M3)), dim3(kT), 0, as_stream(stream), p->amp.as<const float>(),
                       amplitudes_, M3);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_phase_set_amplitudes(cmtip_phaser p, const float* amplitudes_, void* stream) {
    if (!valid(p)) return CMTIP_ERR_ARG;
    if (!amplitudes_) {
        set_error("phase_set_amplitudes: null input");
        return CMTIP_ERR_ARG;
    }
    const int64_t M3 = (int64_t)p->M * p->M * p->M;
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(M3)), dim3(kT), 0, as_stream(stream), amplitudes_, p->amp.as<float>(),
                       M3);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_phase_iterate(cmtip_phaser p, float* rho_, const float* support_, int mode, int n,
                                   double beta, double rho_max, void* stream) {
    if (!valid(p)) return CMTIP_ERR_ARG;
    if (!rho_ || (mode != 0 && !support_) || mode < 0 || mode > 2 || n < 0) {
        set_error("phase_iterate: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipStream_t st = as_stream(stream);
    const float rm = (float)rho_max;  // +inf stays +inf
    for (int k = 0; k < n; ++k) CMTIP_TRY(phase_step(p, rho_, support_, mode, (float)beta, rm, st));
    return CMTIP_OK;
}

extern "C" int cmtip_phase_shrink_wrap(cmtip_phaser p, const float* rho_, double cutoff, double sigma,
                                       float* support_, void* stream) {
    if (!valid(p)) return CMTIP_ERR_ARG;
    if (!rho_ || !support_) {
        set_error("phase_shrink_wrap: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipStream_t st = as_stream(stream);
    const int64_t M3 = (int64_t)p->M * p->M * p->M;
    // rho_gauss_ = gaussian_filter(|rho_|, 'wrap', sigma, truncate 2) (phase.py:158-161)
    CMTIP_TRY(gauss3(p, rho_, 2, sigma, 2.0, 1, st));
    CMTIP_TRY(max_into_slot(p, nullptr, rho_, 1, st));
    hipLaunchKernelGGL(threshold_kernel, dim3(nblk(M3)), dim3(kT), 0, st, p->da.as<const double>(),
                       (const unsigned long long*)max_slot(p), cutoff, support_, M3);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_phase_recenter(cmtip_phaser p, float* rho_, float* support_, int* shift, void* stream) {
    if (!valid(p)) return CMTIP_ERR_ARG;
    if (!rho_ || !support_) {
        set_error("phase_recenter: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipStream_t st = as_stream(stream);
    const int M = p->M;
    const int64_t M3 = (int64_t)M * M * M;
    double* part = p->red.as<double>();
    double* sums = part + 4 * kRed;
    hipLaunchKernelGGL(com_partial_kernel, dim3(kRed), dim3(kT), 0, st, (const float*)rho_, M, part);
    hipLaunchKernelGGL(com_finish_kernel, dim3(1), dim3(64), 0, st, (const double*)part, kRed, sums);
    CMTIP_LAUNCH_CHECK();
    double h[4];
    CMTIP_HIP(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, st));
    CMTIP_HIP(hipStreamSynchronize(st));
    // vect = np.round(num / den * M / 2) (phase.py:23): round half to even
    int s[3];
    for (int d = 0; d < 3; ++d) {
        const double v = nearbyint(h[d + 1] / h[0] * M / 2);
        s[d] = (h[0] > 0.0 && isfinite(v)) ? (int)v : 0;
        if (shift) shift[d] = s[d];
    }
    float* tmp = reinterpret_cast<float*>(p->da.p);
    for (float* a : {rho_, support_}) {
        hipLaunchKernelGGL(roll_kernel, dim3(nblk(M3)), dim3(kT), 0, st, (const float*)a, tmp, M, s[0], s[1], s[2]);
        hipLaunchKernelGGL(copy_kernel, dim3(nblk(M3)), dim3(kT), 0, st, (const float*)tmp, a, M3);
    }
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_phase_ac(cmtip_phaser p, const float* rho_, float* ac_phased, void* stream) {
    if (!valid(p)) return CMTIP_ERR_ARG;
    if (!rho_ || !ac_phased) {
        set_error("phase_ac: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipStream_t st = as_stream(stream);
    const int M = p->M;
    float2* t = p->t.as<float2>();
    // fftshift(|ifftn(|fftn(rho_)|^2)|) (phase.py:214-217); the shift of each axis is
    // applied by the inverse pass of that axis
    CMTIP_TRY(forward_21(p, rho_, st));
    AxisDesc d;
    axis_lines(d, M, 0, true, true);
    d.in = t;
    d.out = t;
    d.sign = -1;
    d.mid_abs2 = 1;
    shift_out(d, M);
    CMTIP_TRY(fft_axis_pass(p->fft, d, st));
    d = AxisDesc();
    axis_lines(d, M, 1, true, true);
    d.in = t;
    d.out = t;
    d.sign = 1;
    shift_out(d, M);
    CMTIP_TRY(fft_axis_pass(p->fft, d, st));
    d = AxisDesc();
    axis_lines(d, M, 2, true, true);
    d.in = t;
    d.out = ac_phased;
    d.out_mode = 3;
    d.out_scale = (float)(1.0 / ((double)M * M * M));
    d.sign = 1;
    shift_out(d, M);
    return fft_axis_pass(p->fft, d, st);
}
