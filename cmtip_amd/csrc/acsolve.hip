// Autocorrelation solve: the W x = d system of cmtip/autocorrelation/autocorrelation.py
// and its scipy CG loop, on the device.
//
// core_problem_convolution (autocorrelation.py:32-63) is a circular convolution on
// an L^3 grid (L even, L >= 2M).  Circular convolution commutes with shifts, so the
// ifftshift of the zero-padded input and the fftshift + crop of the output
// (:49-56) cancel: u is placed at positions [0, M) of every axis and the result is
// read from the same positions.  The five passes are: FFT axis 2 (M^2 lines, M
// nonzero inputs each), FFT axis 1 (M L lines), FFT axis 0 -> multiply by the
// kernel spectrum -> inverse FFT axis 0 in one LDS-resident pass (L^2 lines),
// inverse axis 1, inverse axis 2 with the crop, support, scale and the rho*u term
// folded into the store.  fourier_reg (:88-110) is the same five-pass shape on the
// M^3 grid with the (ifftshifted) antisupport mask as the multiplier.
//
// The CG operator W = alpha conv + rho I + phi reg (autocorrelation.py:150-155) runs
// as ONE five-pass sequence on half the spectrum:
//  - fourier_reg folds into the Toeplitz multiplier when L is a multiple c of M (the
//    reference's M_ups = 2M): the M-point DFT of the input equals the L-point DFT of
//    the zero-padded input at multiples of c, so the antisupport becomes an L^3
//    multiplier that is zero off the multiples of c.
//  - u and the result are real and only the real part is kept (:59-62, :108-109), so
//    a real multiplier G can be replaced by its symmetric part (G[k] + G[-k]) / 2,
//    which makes every intermediate Hermitian: passes 1-4 keep only the first L/2 + 1
//    positions of the last axis and pass 5 extends them by conjugate symmetry.
//  - the CG direction update runs inside pass 1's load and the p.q partials inside
//    pass 5's store; one kernel updates x and r, one block reduces r.r and runs the
//    loop check, so an iteration is seven launches.
#include <math.h>

#include <algorithm>
#include <vector>

#include "fft.h"

using namespace cmtip;

struct cmtip_acsolver_s {
    int M = 0, L = 0;
    Fft1D fftL, fftM;
    DevBuf Fk;       // L^3 float, kernel spectrum in FFT position order
    DevBuf mask;     // M^3 float, ifftshift(F_antisupport) in FFT position order
    DevBuf tmp1;     // max(M^2 L, M^3) float2
    DevBuf tmp2;     // M L^2 float2
    DevBuf r, p, q;  // M^3 float CG vectors
    DevBuf scal;     // FP64 CG scalars
    DevBuf part;     // kRedBlocks doubles: per-block partial sums (fixed-order reductions)
    DevBuf ctrl;     // int32 CG state: stop, converged, iterations, maxiter
    DevBuf Gh;       // L * L * (L/2 + 1) float: folded, symmetrised multiplier of W (half spectrum)
    DevBuf pq_part;  // per-workgroup partial sums of p.q
    double g_alpha = 0.0, g_phi = 0.0;
    bool g_valid = false;  // Gh matches Fk, mask, g_alpha and g_phi
    int* host_state = nullptr;  // pinned host-mapped mirror: stop, converged, iterations, pad, rho (double)
    int* host_state_dev = nullptr;
    const float* support = nullptr;
    const int* cur_skip = nullptr;  // set while CG runs: passes do nothing once it stops
    // CG runs on the solver's own stream (capturable; the legacy default stream is not),
    // ordered after and before the caller's stream with two events
    hipStream_t cg_stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // CG: the iterate and the support mask in solver memory while CG runs, and the
    // instantiated graph of kCgGraphIters iterations with the scalars it was captured for
    DevBuf xbuf, supbuf;
    // set_kernel's two L^3 complex scratch grids, kept so that the call neither allocates
    // nor waits for the device
    DevBuf kscr1, kscr2;
    hipGraphExec_t cg_exec = nullptr;
    double cg_rho = 0.0, cg_phi = 0.0;
    bool cg_with_support = false;
};

namespace {

constexpr int kT = 256;
enum { S_RHO = 0, S_RHO_PREV = 1, S_PQ = 2, S_RR_NEW = 3, S_ATOL = 4, S_B2 = 5, S_COUNT = 8 };

unsigned nblk(int64_t n, int64_t cap = 8192) {
    int64_t b = (n + kT - 1) / kT;
    if (b > cap) b = cap;
    return (unsigned)(b < 1 ? 1 : b);
}

// ---------------------------------------------------------------- gen_F_antisupport
// mask = round(sqrt(H^2 + K^2 + L^2), 4) > pi on np.linspace(-pi, pi, M)^3
// (autocorrelation.py:75-78), in array order.  linspace as numpy computes it:
// i * step - pi, with the last sample set to pi exactly; np.around(x, 4) is
// rint(x * 1e4) / 1e4.
__global__ void antisupport_kernel(float* __restrict__ mask, int M, double step) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
    const int64_t M3 = (int64_t)M * M * M;
    if (t >= M3) return;
    const int i0 = (int)(t / ((int64_t)M * M)), i1 = (int)((t / M) % M), i2 = (int)(t % M);
    auto lin = [&](int i) -> double { return (i == M - 1 && M > 1) ? M_PI : (double)i * step + (-M_PI); };
    const double a = lin(i0), b = lin(i1), c = lin(i2);
    const double q = sqrt(a * a + b * b + c * c);
    const double qr = rint(q * 1e4) / 1e4;
    mask[t] = qr > M_PI ? 1.0f : 0.0f;
}

// out[p0][p1][p2] = in[(p0 + h) % M][(p1 + h) % M][(p2 + h) % M], h = M // 2 (ifftshift)
__global__ void ifftshift3_kernel(const float* __restrict__ in, float* __restrict__ out, int M) {
    const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
    const int64_t M3 = (int64_t)M * M * M;
    if (t >= M3) return;
    const int p0 = (int)(t / ((int64_t)M * M)), p1 = (int)((t / M) % M), p2 = (int)(t % M);
    const int h = M / 2;
    const int i0 = (p0 + h) % M, i1 = (p1 + h) % M, i2 = (p2 + h) % M;
    out[t] = in[((int64_t)i0 * M + i1) * M + i2];
}

// ---------------------------------------------------------------- support filter
// scipy.ndimage.gaussian_filter(ac, 0.5): radius int(4 * 0.5 + 0.5) = 2, weights
// exp(-x^2 / (2 sigma^2)) normalised, mode 'reflect' (edge sample repeated), one axis
// at a time; symmetric weights are applied as w0 x[i] + sum_j wj (x[i-j] + x[i+j]) with
// the farthest pair first, the order of scipy's correlate1d.  round_f32 rounds after each axis (float32 input).
__device__ __forceinline__ int reflect(int i, int M) {
    // half-sample symmetric reflection, valid for |overhang| <= M
    if (i < 0) i = -i - 1;
    if (i >= M) i = 2 * M - i - 1;
    return i < 0 ? 0 : (i >= M ? M - 1 : i);
}

__global__ void gauss_axis_kernel(const double* __restrict__ in, double* __restrict__ out, int M, int axis,
                                  double w0, double w1, double w2, int round_f32) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
    const int64_t M3 = (int64_t)M * M * M;
    if (t >= M3) return;
    int idx[3] = {(int)(t / ((int64_t)M * M)), (int)((t / M) % M), (int)(t % M)};
    const int64_t stride = axis == 0 ? (int64_t)M * M : (axis == 1 ? M : 1);
    const int i = idx[axis];
    const int64_t base = t - (int64_t)i * stride;
    auto at = [&](int j) -> double { return in[base + (int64_t)reflect(j, M) * stride]; };
    double v = at(i) * w0;
    v += (at(i - 2) + at(i + 2)) * w2;
    v += (at(i - 1) + at(i + 1)) * w1;
    if (round_f32) v = (double)(float)v;
    out[t] = v;
}

__global__ void threshold_kernel(const double* __restrict__ in, float* __restrict__ out, int64_t n, double thr) {
    const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (t < n) out[t] = in[t] > thr ? 1.0f : 0.0f;
}

// ---------------------------------------------------------------- CG vector kernels
//
// scipy.sparse.linalg.cg (iterative.py:375-422) with the loop state on the device:
// ctrl = {stop, converged, iterations, maxiter}.  Every kernel of an iteration does
// nothing once stop is set, so several iterations can be launched (or replayed from
// a captured graph) before the host looks at the state, with exactly scipy's
// iteration count: at the start of iteration i the loop ends when i == maxiter (info
// maxiter) or when ||r|| < atol (info 0).  Reductions are two-stage: per-block
// partial sums, then one block adds them in a fixed order, so results do not depend
// on scheduling.

constexpr int kRedBlocks = 512;
enum { C_STOP = 0, C_CONV = 1, C_ITERS = 2, C_MAXITER = 3 };

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ void block_partial(double v, double* __restrict__ part) {
    __shared__ double red[kT / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kT / 64; ++i) t += red[i];
        part[blockIdx.x] = t;
    }
}

__device__ __forceinline__ bool stopped(const int* ctrl) { return ctrl && ctrl[C_STOP]; }

__global__ void dot_partial_kernel(const float* __restrict__ a, const float* __restrict__ b, int64_t n,
                                   double* __restrict__ part, const int* __restrict__ ctrl) {
    if (stopped(ctrl)) return;
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kT;
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += stride) s += (double)a[i] * (double)b[i];
    block_partial(s, part);
}

// s[slot] = sum of the kRedBlocks partials (fixed order)
__global__ void reduce_partials_kernel(const double* __restrict__ part, double* __restrict__ s, int slot,
                                       const int* __restrict__ ctrl) {
    if (stopped(ctrl)) return;
    __shared__ double red[kRedBlocks];
    red[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int h = kRedBlocks / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) s[slot] = red[0];
}

__global__ void set_scalars_kernel(double* __restrict__ s, int first, int count, double v) {
    if (threadIdx.x < count) s[first + threadIdx.x] = v;
}

// CG state init: stop when b == 0 (scipy returns b itself), atol = rtol ||b||
__global__ void cg_init_kernel(double* __restrict__ s, int* __restrict__ ctrl, int maxiter, double rtol) {
    if (threadIdx.x != 0) return;
    const double b2 = s[S_B2];
    s[S_ATOL] = rtol * sqrt(b2);
    ctrl[C_STOP] = b2 == 0.0 ? 1 : 0;
    ctrl[C_CONV] = b2 == 0.0 ? 1 : 0;
    ctrl[C_ITERS] = 0;
    ctrl[C_MAXITER] = maxiter;
}

// start of an iteration: loop exhausted or converged?  else count it
__global__ void cg_check_kernel(const double* __restrict__ s, int* __restrict__ ctrl) {
    if (threadIdx.x != 0 || ctrl[C_STOP]) return;
    if (ctrl[C_ITERS] >= ctrl[C_MAXITER]) {
        ctrl[C_STOP] = 1;
        return;
    }
    if (sqrt(s[S_RHO]) < s[S_ATOL]) {
        ctrl[C_CONV] = 1;
        ctrl[C_STOP] = 1;
        return;
    }
    ctrl[C_ITERS] += 1;
}

// r = b - Ax
__global__ void residual_kernel(const float* __restrict__ b, const float* __restrict__ ax, float* __restrict__ r,
                                int64_t n, const int* __restrict__ ctrl) {
    if (stopped(ctrl)) return;
    const int64_t stride = (int64_t)gridDim.x * kT;
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += stride) r[i] = b[i] - ax[i];
}

// x = b when b == 0 (the early return of scipy's cg)
__global__ void copy_if_zero_rhs_kernel(const float* __restrict__ b, float* __restrict__ x, int64_t n,
                                        const double* __restrict__ s) {
    if (s[S_B2] != 0.0) return;
    const int64_t stride = (int64_t)gridDim.x * kT;
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += stride) x[i] = b[i];
}

// alpha = rho / (p.q), with p.q the fixed-order sum of the npq partials (every
// workgroup adds them in the same order, so all agree bit for bit); x += alpha p;
// r -= alpha q; per-workgroup partials of r.r.
__global__ void update_xr_kernel(float* __restrict__ x, float* __restrict__ r, const float* __restrict__ p,
                                 const float* __restrict__ q, const double* __restrict__ s,
                                 const double* __restrict__ pq_part, int npq, int64_t n,
                                 double* __restrict__ part, const int* __restrict__ ctrl) {
    if (ctrl[C_STOP]) return;
    __shared__ double red[kT / 64];
    __shared__ double pq_s;
    double v = 0.0;
    for (int i = threadIdx.x; i < npq; i += kT) v += pq_part[i];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kT / 64; ++i) t += red[i];
        pq_s = t;
    }
    __syncthreads();
    const float alpha = (float)(s[S_RHO] / pq_s);
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kT;
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += stride) {
        x[i] = x[i] + alpha * p[i];
        const float rn = r[i] - alpha * q[i];
        r[i] = rn;
        acc += (double)rn * (double)rn;
    }
    block_partial(acc, part);
}

// rho_prev = rho, rho = sum of the kRedBlocks r.r partials (fixed order), then the loop
// check of the next iteration (cg_check_kernel's rule)
__global__ void rotate_check_kernel(const double* __restrict__ part, double* __restrict__ s, int* __restrict__ ctrl) {
    if (ctrl[C_STOP]) return;
    __shared__ double red[kRedBlocks];
    red[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int h = kRedBlocks / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double rho = red[0];
    s[S_RHO_PREV] = s[S_RHO];
    s[S_RHO] = rho;
    if (ctrl[C_ITERS] >= ctrl[C_MAXITER]) {
        ctrl[C_STOP] = 1;
    } else if (sqrt(rho) < s[S_ATOL]) {
        ctrl[C_CONV] = 1;
        ctrl[C_STOP] = 1;
    } else {
        ctrl[C_ITERS] += 1;
    }
}

// The operator's multiplier on the half spectrum (p0, p1 < L, p2 <= L/2):
// G = ka (K[p] + K[-p]) / 2 + ma (m[p / c] + m[-p / c]) / 2 where the second term is
// present only at positions whose three indices are multiples of c (c = 0: absent).
// K is the kernel spectrum (L^3), m the antisupport in FFT order (M^3); indices are
// taken modulo the grid edge.
__global__ void half_multiplier_kernel(const float* __restrict__ K, const float* __restrict__ m,
                                       float* __restrict__ G, int L, int M, int c, double ka, double ma) {
    const int64_t H = L / 2 + 1;
    const int64_t total = (int64_t)L * L * H;
    const int64_t stride = (int64_t)gridDim.x * kT;
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < total; i += stride) {
        const int p2 = (int)(i % H);
        const int64_t t = i / H;
        const int p1 = (int)(t % L);
        const int p0 = (int)(t / L);
        const int q0 = p0 ? L - p0 : 0, q1 = p1 ? L - p1 : 0, q2 = p2 ? L - p2 : 0;
        const double k = 0.5 * ((double)K[((int64_t)p0 * L + p1) * L + p2] + (double)K[((int64_t)q0 * L + q1) * L + q2]);
        double g = ka * k;
        if (c > 0 && p0 % c == 0 && p1 % c == 0 && p2 % c == 0) {
            const int j0 = p0 / c, j1 = p1 / c, j2 = p2 / c;
            const int k0 = j0 ? M - j0 : 0, k1 = j1 ? M - j1 : 0, k2 = j2 ? M - j2 : 0;
            g += ma * 0.5 * ((double)m[((int64_t)j0 * M + j1) * M + j2] + (double)m[((int64_t)k0 * M + k1) * M + k2]);
        }
        G[i] = (float)g;
    }
}

// copy the loop state to pinned host memory (read by the host after a sync)
__global__ void publish_kernel(const int* __restrict__ ctrl, const double* __restrict__ s, int* __restrict__ host) {
    if (threadIdx.x != 0) return;
    host[0] = ctrl[C_STOP];
    host[1] = ctrl[C_CONV];
    host[2] = ctrl[C_ITERS];
    *reinterpret_cast<double*>(host + 4) = s[S_RHO];
}

bool valid(cmtip_acsolver s) {
    if (!s) {
        set_error("null solver");
        return false;
    }
    return true;
}

// out = out_scale * S (.) conv(S (.) u) + add_scale * add_src   (add_src may be null)
int conv_apply(cmtip_acsolver_s* s, const float* u, float* out, float out_scale, const float* add_src,
               float add_scale, hipStream_t st) {
    const int M = s->M, L = s->L;
    const int64_t LL = (int64_t)L * L;
    float2* t1 = s->tmp1.as<float2>();
    float2* t2 = s->tmp2.as<float2>();
    AxisDesc d;
    // 1: axis 2, lines (i0, i1), input positions [0, M)
    d = AxisDesc();
    d.in = u;
    d.in_real = 1;
    d.in_su = (int64_t)M * M;
    d.in_sv = M;
    d.in_sa = 1;
    d.ia = M;
    d.ig1 = 0;
    d.in_mask = s->support;
    d.out = t1;
    d.out_su = (int64_t)M * L;
    d.out_sv = L;
    d.out_sa = 1;
    d.oa = L;
    d.nu = M;
    d.nv = M;
    d.sign = -1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftL, d, st));
    // 2: axis 1, lines (i0, p2)
    d = AxisDesc();
    d.in = t1;
    d.in_su = (int64_t)M * L;
    d.in_sv = 1;
    d.in_sa = L;
    d.ia = M;
    d.ig1 = 0;
    d.out = t2;
    d.out_su = LL;
    d.out_sv = 1;
    d.out_sa = L;
    d.oa = L;
    d.nu = M;
    d.nv = L;
    d.sign = -1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftL, d, st));
    // 3: axis 0 forward, x kernel spectrum, inverse; lines (p1, p2); in place
    d = AxisDesc();
    d.in = t2;
    d.in_su = L;
    d.in_sv = 1;
    d.in_sa = LL;
    d.ia = M;
    d.ig1 = 0;
    d.out = t2;
    d.out_su = L;
    d.out_sv = 1;
    d.out_sa = LL;
    d.oa = M;
    d.og1 = 0;
    d.nu = L;
    d.nv = L;
    d.sign = -1;
    d.mid_r = s->Fk.as<const float>();
    d.mid_su = L;
    d.mid_sv = 1;
    d.mid_sa = LL;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftL, d, st));
    // 4: inverse axis 1, lines (i0, p2), keep positions [0, M)
    d = AxisDesc();
    d.in = t2;
    d.in_su = LL;
    d.in_sv = 1;
    d.in_sa = L;
    d.ia = L;
    d.out = t1;
    d.out_su = (int64_t)M * L;
    d.out_sv = 1;
    d.out_sa = L;
    d.oa = M;
    d.nu = M;
    d.nv = L;
    d.sign = +1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftL, d, st));
    // 5: inverse axis 2, lines (i0, i1), real part, crop, support, scale, axpy
    d = AxisDesc();
    d.in = t1;
    d.in_su = (int64_t)M * L;
    d.in_sv = L;
    d.in_sa = 1;
    d.ia = L;
    d.out = out;
    d.out_mode = 1;
    d.out_su = (int64_t)M * M;
    d.out_sv = M;
    d.out_sa = 1;
    d.oa = M;
    d.out_scale = out_scale;
    d.out_mask = s->support;
    d.add_src = add_src;
    d.add_scale = add_scale;
    d.nu = M;
    d.nv = M;
    d.sign = +1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftL, d, st));
    return CMTIP_OK;
}

// out = out_scale * S (.) fftshift(ifftn(fftn(ifftshift(S (.) u)) * mask)) + add_scale * add_src.
// By the shift theorem the ifftshift / fftshift pair cancels once the mask is held in
// FFT position order (np.fft.ifftshift(F_antisupport), autocorrelation.py:105), so u
// is transformed in place and the result read back at the same positions.
int reg_apply(cmtip_acsolver_s* s, const float* u, float* out, float out_scale, const float* add_src,
              float add_scale, hipStream_t st) {
    const int M = s->M;
    const int64_t MM = (int64_t)M * M;
    float2* t1 = s->tmp1.as<float2>();
    float2* t2 = s->tmp2.as<float2>();
    AxisDesc d;
    // 1: axis 2
    d = AxisDesc();
    d.in = u;
    d.in_real = 1;
    d.in_su = MM;
    d.in_sv = M;
    d.in_sa = 1;
    d.ia = M;
    d.in_mask = s->support;
    d.out = t1;
    d.out_su = MM;
    d.out_sv = M;
    d.out_sa = 1;
    d.oa = M;
    d.nu = M;
    d.nv = M;
    d.sign = -1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftM, d, st));
    // 2: axis 1
    d = AxisDesc();
    d.in = t1;
    d.in_su = MM;
    d.in_sv = 1;
    d.in_sa = M;
    d.ia = M;
    d.out = t2;
    d.out_su = MM;
    d.out_sv = 1;
    d.out_sa = M;
    d.oa = M;
    d.nu = M;
    d.nv = M;
    d.sign = -1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftM, d, st));
    // 3: axis 0 forward, x mask, inverse (in place)
    d = AxisDesc();
    d.in = t2;
    d.in_su = M;
    d.in_sv = 1;
    d.in_sa = MM;
    d.ia = M;
    d.out = t2;
    d.out_su = M;
    d.out_sv = 1;
    d.out_sa = MM;
    d.oa = M;
    d.nu = M;
    d.nv = M;
    d.sign = -1;
    d.mid_r = s->mask.as<const float>();
    d.mid_su = M;
    d.mid_sv = 1;
    d.mid_sa = MM;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftM, d, st));
    // 4: inverse axis 1
    d = AxisDesc();
    d.in = t2;
    d.in_su = MM;
    d.in_sv = 1;
    d.in_sa = M;
    d.ia = M;
    d.out = t1;
    d.out_su = MM;
    d.out_sv = 1;
    d.out_sa = M;
    d.oa = M;
    d.nu = M;
    d.nv = M;
    d.sign = +1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftM, d, st));
    // 5: inverse axis 2, real part, support, scale, axpy
    d = AxisDesc();
    d.in = t1;
    d.in_su = MM;
    d.in_sv = M;
    d.in_sa = 1;
    d.ia = M;
    d.out = out;
    d.out_mode = 1;
    d.out_su = MM;
    d.out_sv = M;
    d.out_sa = 1;
    d.oa = M;
    d.out_scale = out_scale;
    d.out_mask = s->support;
    d.add_src = add_src;
    d.add_scale = add_scale;
    d.nu = M;
    d.nv = M;
    d.sign = +1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftM, d, st));
    return CMTIP_OK;
}

// c = L / M when fourier_reg folds into the L-grid multiplier, else 0
int fold_factor(const cmtip_acsolver_s* s) { return s->L % s->M == 0 ? s->L / s->M : 0; }

// Gh for W = alpha conv + rho I + phi reg (rebuilt when alpha, phi, the spectrum or the
// antisupport changed)
int ensure_multiplier(cmtip_acsolver_s* s, double alpha, double phi, hipStream_t st) {
    if (s->g_valid && s->g_alpha == alpha && s->g_phi == phi) return CMTIP_OK;
    const int M = s->M, L = s->L, c = fold_factor(s);
    const double L3 = (double)L * L * L, M3 = (double)M * M * M;
    const int64_t total = (int64_t)L * L * (L / 2 + 1);
    hipLaunchKernelGGL(half_multiplier_kernel, dim3(nblk(total, 4096)), dim3(kT), 0, st, s->Fk.as<const float>(),
                       s->mask.as<const float>(), s->Gh.as<float>(), L, M, c, alpha / L3, c ? phi / M3 : 0.0);
    CMTIP_LAUNCH_CHECK();
    s->g_alpha = alpha;
    s->g_phi = phi;
    s->g_valid = true;
    return CMTIP_OK;
}

// out = S (.) Re(ifft(Gh fft(S (.) u))) + rho u in five passes over half the spectrum
// of the last axis (H = L/2 + 1 positions): t1 is (M, M, H), t2 is (M, L, H).  With
// cg set, u is r: pass 1 first replaces p by the new CG direction and transforms that,
// pass 5 adds rho p and, when pq_parts is non-null, writes the p.q partials to pq_part
// (their count to *pq_parts).
int fast_apply(cmtip_acsolver_s* s, const float* u, float* out, float rho, bool cg, int* pq_parts,
               hipStream_t st) {
    const int M = s->M, L = s->L, H = L / 2 + 1;
    const int64_t MM = (int64_t)M * M, MH = (int64_t)M * H, LH = (int64_t)L * H;
    float2* t1 = s->tmp1.as<float2>();
    float2* t2 = s->tmp2.as<float2>();
    float* p = s->p.as<float>();
    AxisDesc d;
    // 1: axis 2 forward, lines (i0, i1), inputs [0, M), outputs [0, H)
    d = AxisDesc();
    d.in = u;
    d.in_real = 1;
    d.in_su = MM;
    d.in_sv = M;
    d.in_sa = 1;
    d.ia = M;
    d.in_mask = s->support;
    if (cg) {
        d.cg_p = p;
        d.cg_scal = s->scal.as<const double>() + S_RHO;  // rho, rho_prev
        d.cg_ctrl = s->ctrl.as<const int>();
    }
    d.out = t1;
    d.out_su = MH;
    d.out_sv = H;
    d.out_sa = 1;
    d.oa = H;
    d.nu = M;
    d.nv = M;
    d.sign = -1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftL, d, st));
    // 2: axis 1 forward, lines (i0, p2)
    d = AxisDesc();
    d.in = t1;
    d.in_su = MH;
    d.in_sv = 1;
    d.in_sa = H;
    d.ia = M;
    d.out = t2;
    d.out_su = LH;
    d.out_sv = 1;
    d.out_sa = H;
    d.oa = L;
    d.nu = M;
    d.nv = H;
    d.sign = -1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftL, d, st));
    // 3: axis 0 forward, x Gh, inverse; lines (p1, p2); in place, outputs [0, M)
    d = AxisDesc();
    d.in = t2;
    d.in_su = H;
    d.in_sv = 1;
    d.in_sa = LH;
    d.ia = M;
    d.out = t2;
    d.out_su = H;
    d.out_sv = 1;
    d.out_sa = LH;
    d.oa = M;
    d.mid_r = s->Gh.as<const float>();
    d.mid_su = H;
    d.mid_sv = 1;
    d.mid_sa = LH;
    d.nu = L;
    d.nv = H;
    d.sign = -1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftL, d, st));
    // 4: axis 1 inverse, lines (i0, p2), outputs [0, M)
    d = AxisDesc();
    d.in = t2;
    d.in_su = LH;
    d.in_sv = 1;
    d.in_sa = H;
    d.ia = L;
    d.out = t1;
    d.out_su = MH;
    d.out_sv = 1;
    d.out_sa = H;
    d.oa = M;
    d.nu = M;
    d.nv = H;
    d.sign = +1;
    d.skip = s->cur_skip;
    CMTIP_TRY(fft_axis_pass(s->fftL, d, st));
    // 5: axis 2 inverse from the half spectrum (conjugate symmetry), real part, crop,
    // support, + rho u
    const float* src = cg ? p : u;
    d = AxisDesc();
    d.in = t1;
    d.in_su = MH;
    d.in_sv = H;
    d.in_sa = 1;
    d.ia = H;
    d.ib = L - H;
    d.in_herm = 1;
    d.out = out;
    d.out_mode = 1;
    d.out_su = MM;
    d.out_sv = M;
    d.out_sa = 1;
    d.oa = M;
    d.out_mask = s->support;
    d.add_src = src;
    d.add_scale = rho;
    d.nu = M;
    d.nv = M;
    d.sign = +1;
    d.skip = s->cur_skip;
    if (pq_parts) {
        d.dot_src = src;
        d.dot_part = s->pq_part.as<double>();
        *pq_parts = (int)fft_axis_grid(s->fftL, d);
    }
    CMTIP_TRY(fft_axis_pass(s->fftL, d, st));
    return CMTIP_OK;
}

int matvec(cmtip_acsolver_s* s, const float* u, float* out, double alpha, double rho, double phi,
           hipStream_t st) {
    CMTIP_TRY(ensure_multiplier(s, alpha, phi, st));
    CMTIP_TRY(fast_apply(s, u, out, (float)rho, false, nullptr, st));
    if (!fold_factor(s)) {
        const double M3 = (double)s->M * s->M * s->M;
        CMTIP_TRY(reg_apply(s, u, out, (float)(phi / M3), out, 1.0f, st));
    }
    return CMTIP_OK;
}

// out = a.b in FP64 with a fixed summation order
int dot(cmtip_acsolver_s* sv, const float* a, const float* b, int64_t n, double* scal, int slot, hipStream_t st,
        const int* ctrl) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(kRedBlocks), dim3(kT), 0, st, a, b, n, sv->part.as<double>(), ctrl);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kRedBlocks), 0, st, sv->part.as<const double>(), scal,
                       slot, ctrl);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

int set_scalars(double* s, int first, int count, double v, hipStream_t st) {
    hipLaunchKernelGGL(set_scalars_kernel, dim3(1), dim3(64), 0, st, s, first, count, v);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

}  // namespace

extern "C" int cmtip_acsolver_create(int M, int L, cmtip_acsolver* out) {
    if (!out) {
        set_error("acsolver_create: null output");
        return CMTIP_ERR_ARG;
    }
    *out = nullptr;
    if (M < 1 || M > 1024 || L < 2 * M || (L % 2) != 0 || L > 4096) {
        set_error("acsolver_create: need M >= 1 and an even L >= 2M (M=%d, L=%d)", M, L);
        return CMTIP_ERR_ARG;
    }
    auto* s = new cmtip_acsolver_s();
    s->M = M;
    s->L = L;
    const int64_t M3 = (int64_t)M * M * M;
    int st = fft1d_init(s->fftL, L);
    if (st == CMTIP_OK) st = fft1d_init(s->fftM, M);
    if (st == CMTIP_OK) st = s->Fk.ensure(sizeof(float) * (size_t)L * L * L);
    if (st == CMTIP_OK) st = s->mask.ensure(sizeof(float) * (size_t)M3);
    const int64_t t1 = (int64_t)M * M * L > M3 ? (int64_t)M * M * L : M3;
    if (st == CMTIP_OK) st = s->tmp1.ensure(sizeof(float2) * (size_t)t1);
    if (st == CMTIP_OK) st = s->tmp2.ensure(sizeof(float2) * (size_t)M * L * L);
    if (st == CMTIP_OK) st = s->r.ensure(sizeof(float) * (size_t)M3);
    if (st == CMTIP_OK) st = s->p.ensure(sizeof(float) * (size_t)M3);
    if (st == CMTIP_OK) st = s->q.ensure(sizeof(float) * (size_t)M3);
    if (st == CMTIP_OK) st = s->scal.ensure(sizeof(double) * S_COUNT);
    if (st == CMTIP_OK) st = s->part.ensure(sizeof(double) * kRedBlocks);
    if (st == CMTIP_OK) st = s->ctrl.ensure(sizeof(int) * 4);
    if (st == CMTIP_OK) st = s->Gh.ensure(sizeof(float) * (size_t)L * L * (L / 2 + 1));
    if (st == CMTIP_OK) st = s->pq_part.ensure(sizeof(double) * (size_t)(M * M > kRedBlocks ? M * M : kRedBlocks));
    if (st == CMTIP_OK) {
        void* h = nullptr;
        if (hipHostMalloc(&h, 64, hipHostMallocMapped) != hipSuccess ||
            hipHostGetDevicePointer((void**)&s->host_state_dev, h, 0) != hipSuccess) {
            set_error("acsolver_create: pinned host state allocation failed");
            st = CMTIP_ERR_ALLOC;
        }
        s->host_state = (int*)h;
    }
    if (st == CMTIP_OK && (hipStreamCreateWithFlags(&s->cg_stream, hipStreamNonBlocking) != hipSuccess ||
                           hipEventCreateWithFlags(&s->ev_in, hipEventDisableTiming) != hipSuccess ||
                           hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming) != hipSuccess)) {
        set_error("acsolver_create: stream / event creation failed");
        st = CMTIP_ERR_HIP;
    }
    if (st == CMTIP_OK) {
        DevBuf arr;
        st = arr.ensure(sizeof(float) * (size_t)M3);
        if (st == CMTIP_OK) st = cmtip_gen_F_antisupport(M, arr.as<float>(), nullptr);
        if (st == CMTIP_OK) st = cmtip_acsolver_set_antisupport(s, arr.as<const float>(), nullptr);
        if (st == CMTIP_OK && hipDeviceSynchronize() != hipSuccess) {
            set_error("acsolver_create: antisupport kernels failed");
            st = CMTIP_ERR_LAUNCH;
        }
        arr.release();
    }
    if (st != CMTIP_OK) {
        cmtip_acsolver_destroy(s);
        return st;
    }
    *out = s;
    return CMTIP_OK;
}

extern "C" int cmtip_gen_F_antisupport(int M, float* out, void* stream) {
    if (M < 1 || !out) {
        set_error("gen_F_antisupport: bad arguments");
        return CMTIP_ERR_ARG;
    }
    const int64_t M3 = (int64_t)M * M * M;
    const double step = M > 1 ? (M_PI - (-M_PI)) / (double)(M - 1) : 0.0;
    hipLaunchKernelGGL(antisupport_kernel, dim3((unsigned)((M3 + kT - 1) / kT)), dim3(kT), 0, as_stream(stream),
                       out, M, step);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

extern "C" int cmtip_acsolver_set_antisupport(cmtip_acsolver s, const float* F_antisupport, void* stream) {
    if (!valid(s)) return CMTIP_ERR_ARG;
    if (!F_antisupport) {
        set_error("acsolver_set_antisupport: null mask");
        return CMTIP_ERR_ARG;
    }
    const int M = s->M;
    const int64_t M3 = (int64_t)M * M * M;
    hipLaunchKernelGGL(ifftshift3_kernel, dim3((unsigned)((M3 + kT - 1) / kT)), dim3(kT), 0, as_stream(stream),
                       F_antisupport, s->mask.as<float>(), M);
    CMTIP_LAUNCH_CHECK();
    s->g_valid = false;
    return CMTIP_OK;
}

extern "C" int cmtip_acsolver_set_spectrum(cmtip_acsolver s, const float* F_real, void* stream) {
    if (!valid(s)) return CMTIP_ERR_ARG;
    if (!F_real) {
        set_error("acsolver_set_spectrum: null spectrum");
        return CMTIP_ERR_ARG;
    }
    const size_t bytes = sizeof(float) * (size_t)s->L * s->L * s->L;
    CMTIP_HIP(hipMemcpyAsync(s->Fk.p, F_real, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    s->g_valid = false;
    return CMTIP_OK;
}

extern "C" int cmtip_acsolver_destroy(cmtip_acsolver s) {
    if (!s) return CMTIP_OK;
    fft1d_free(s->fftL);
    fft1d_free(s->fftM);
    s->Fk.release();
    s->mask.release();
    s->tmp1.release();
    s->tmp2.release();
    s->r.release();
    s->p.release();
    s->q.release();
    s->scal.release();
    s->part.release();
    s->ctrl.release();
    s->Gh.release();
    s->pq_part.release();
    if (s->cg_exec) (void)hipGraphExecDestroy(s->cg_exec);
    s->xbuf.release();
    s->supbuf.release();
    s->kscr1.release();
    s->kscr2.release();
    if (s->host_state) (void)hipHostFree(s->host_state);
    if (s->ev_in) (void)hipEventDestroy(s->ev_in);
    if (s->ev_out) (void)hipEventDestroy(s->ev_out);
    if (s->cg_stream) (void)hipStreamDestroy(s->cg_stream);
    delete s;
    return CMTIP_OK;
}

extern "C" int cmtip_ac_support(const double* ac_estimate, int M, float* support, void* stream) {
    return cmtip_ac_support_ex(ac_estimate, M, 0, support, stream);
}

extern "C" int cmtip_ac_support_ex(const double* ac_estimate, int M, int round_f32, float* support,
                                   void* stream) {
    if (!ac_estimate || !support || M < 1) {
        set_error("ac_support: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipStream_t st = as_stream(stream);
    const int64_t M3 = (int64_t)M * M * M;
    // weights in numpy's order: x = arange(-2, 3); phi = exp(-0.5 / 0.25 * x**2); phi / phi.sum()
    double phi[5];
    for (int i = 0; i < 5; ++i) {
        const double x = (double)(i - 2);
        phi[i] = exp(-0.5 / 0.25 * (x * x));
    }
    double sum = 0.0;
    for (int i = 0; i < 5; ++i) sum += phi[i];
    const double w0 = phi[2] / sum, w1 = phi[3] / sum, w2 = phi[4] / sum;
    DevBuf a, b;
    CMTIP_TRY(a.ensure(sizeof(double) * (size_t)M3));
    CMTIP_TRY(b.ensure(sizeof(double) * (size_t)M3));
    const unsigned g = (unsigned)((M3 + kT - 1) / kT);
    hipLaunchKernelGGL(gauss_axis_kernel, dim3(g), dim3(kT), 0, st, ac_estimate, a.as<double>(), M, 0, w0, w1, w2,
                       round_f32);
    hipLaunchKernelGGL(gauss_axis_kernel, dim3(g), dim3(kT), 0, st, a.as<const double>(), b.as<double>(), M, 1, w0,
                       w1, w2, round_f32);
    hipLaunchKernelGGL(gauss_axis_kernel, dim3(g), dim3(kT), 0, st, b.as<const double>(), a.as<double>(), M, 2, w0,
                       w1, w2, round_f32);
    hipLaunchKernelGGL(threshold_kernel, dim3(g), dim3(kT), 0, st, a.as<const double>(), support, M3, 1e-12);
    hipError_t e = hipGetLastError();
    hipError_t e2 = hipStreamSynchronize(st);  // scratch buffers are freed below
    a.release();
    b.release();
    if (e != hipSuccess || e2 != hipSuccess) {
        set_error("ac_support: %s", hipGetErrorString(e != hipSuccess ? e : e2));
        return CMTIP_ERR_LAUNCH;
    }
    return CMTIP_OK;
}

extern "C" int cmtip_acsolver_set_support(cmtip_acsolver s, const float* support) {
    if (!valid(s)) return CMTIP_ERR_ARG;
    s->support = support;
    return CMTIP_OK;
}

extern "C" int cmtip_acsolver_set_kernel(cmtip_acsolver s, const float* ugrid_conv, void* stream) {
    if (!valid(s)) return CMTIP_ERR_ARG;
    if (!ugrid_conv) {
        set_error("acsolver_set_kernel: null kernel");
        return CMTIP_ERR_ARG;
    }
    hipStream_t st = as_stream(stream);
    const int L = s->L;
    const int64_t LL = (int64_t)L * L;
    CMTIP_TRY(s->kscr1.ensure(sizeof(float2) * (size_t)LL * L));
    CMTIP_TRY(s->kscr2.ensure(sizeof(float2) * (size_t)LL * L));
    DevBuf& big1 = s->kscr1;
    DevBuf& big2 = s->kscr2;
    const int h = L / 2;
    AxisDesc d;
    // axis 2 with ifftshift
    d = AxisDesc();
    d.in = ugrid_conv;
    d.in_real = 1;
    d.in_su = LL;
    d.in_sv = L;
    d.in_sa = 1;
    d.ia = L - h;
    d.ig1 = h;
    d.ib = h;
    d.ig2 = 0;
    d.out = big1.p;
    d.out_su = LL;
    d.out_sv = L;
    d.out_sa = 1;
    d.oa = L;
    d.nu = L;
    d.nv = L;
    d.sign = -1;
    int rc = fft_axis_pass(s->fftL, d, st);
    // axis 1 with ifftshift
    if (rc == CMTIP_OK) {
        d = AxisDesc();
        d.in = big1.p;
        d.in_su = LL;
        d.in_sv = 1;
        d.in_sa = L;
        d.ia = L - h;
        d.ig1 = h;
        d.ib = h;
        d.ig2 = 0;
        d.out = big2.p;
        d.out_su = LL;
        d.out_sv = 1;
        d.out_sa = L;
        d.oa = L;
        d.nu = L;
        d.nv = L;
        d.sign = -1;
        rc = fft_axis_pass(s->fftL, d, st);
    }
    // axis 0 with ifftshift, real part out
    if (rc == CMTIP_OK) {
        d = AxisDesc();
        d.in = big2.p;
        d.in_su = L;
        d.in_sv = 1;
        d.in_sa = LL;
        d.ia = L - h;
        d.ig1 = h;
        d.ib = h;
        d.ig2 = 0;
        d.out = s->Fk.p;
        d.out_mode = 1;
        d.out_su = L;
        d.out_sv = 1;
        d.out_sa = LL;
        d.oa = L;
        d.nu = L;
        d.nv = L;
        d.sign = -1;
        rc = fft_axis_pass(s->fftL, d, st);
    }
    s->g_valid = false;
    return rc;
}

extern "C" int cmtip_core_problem_convolution(cmtip_acsolver s, const float* u, float* out, void* stream) {
    if (!valid(s)) return CMTIP_ERR_ARG;
    if (!u || !out || u == out) {
        set_error("core_problem_convolution: need distinct u and out");
        return CMTIP_ERR_ARG;
    }
    const double L3 = (double)s->L * s->L * s->L;
    return conv_apply(s, u, out, (float)(1.0 / L3), nullptr, 0.0f, as_stream(stream));
}

extern "C" int cmtip_fourier_reg(cmtip_acsolver s, const float* u, float* out, void* stream) {
    if (!valid(s)) return CMTIP_ERR_ARG;
    if (!u || !out || u == out) {
        set_error("fourier_reg: need distinct u and out");
        return CMTIP_ERR_ARG;
    }
    const double M3 = (double)s->M * s->M * s->M;
    return reg_apply(s, u, out, (float)(1.0 / M3), nullptr, 0.0f, as_stream(stream));
}

extern "C" int cmtip_acsolver_matvec(cmtip_acsolver s, const float* u, float* out, double alpha, double rho,
                                     double phi, void* stream) {
    if (!valid(s)) return CMTIP_ERR_ARG;
    if (!u || !out || u == out) {
        set_error("acsolver_matvec: need distinct u and out");
        return CMTIP_ERR_ARG;
    }
    return matvec(s, u, out, alpha, rho, phi, as_stream(stream));
}

// One CG iteration, the loop check of the next one at its end (every kernel a no-op
// once ctrl[C_STOP] is set): p = r + beta p and q = W p (five passes, p.q partials in
// the last), then x and r, then rho and the check.  Without the fold the M-grid
// fourier_reg passes and a separate p.q reduction follow the five passes.
static int cg_iteration(cmtip_acsolver_s* s, float* x, double rho, double phi, hipStream_t st) {
    const int64_t n = (int64_t)s->M * s->M * s->M;
    int* ctrl = s->ctrl.as<int>();
    float* r = s->r.as<float>();
    float* p = s->p.as<float>();
    float* q = s->q.as<float>();
    int npq = 0;
    const bool fold = fold_factor(s) > 0;
    CMTIP_TRY(fast_apply(s, r, q, (float)rho, true, fold ? &npq : nullptr, st));
    if (!fold) {
        const double M3 = (double)s->M * s->M * s->M;
        CMTIP_TRY(reg_apply(s, p, q, (float)(phi / M3), q, 1.0f, st));
        hipLaunchKernelGGL(dot_partial_kernel, dim3(kRedBlocks), dim3(kT), 0, st, (const float*)p, (const float*)q, n,
                           s->pq_part.as<double>(), (const int*)ctrl);
        npq = kRedBlocks;
    }
    hipLaunchKernelGGL(update_xr_kernel, dim3(kRedBlocks), dim3(kT), 0, st, x, r, (const float*)p, (const float*)q,
                       (const double*)s->scal.as<double>(), s->pq_part.as<const double>(), npq, n, s->part.as<double>(),
                       (const int*)ctrl);
    hipLaunchKernelGGL(rotate_check_kernel, dim3(1), dim3(kRedBlocks), 0, st, s->part.as<const double>(),
                       s->scal.as<double>(), ctrl);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

constexpr int kCgGraphIters = 5;

static bool use_graphs() {
    const char* e = getenv("CMTIP_CG_GRAPH");
    return !(e && e[0] == '0');
}

namespace {
// s->support points at a solver-owned copy of the caller's mask while CG runs (so that
// the cached iteration graph refers only to solver memory) and back at the caller's
// mask afterwards, on every return path
struct SupportSwap {
    cmtip_acsolver_s* s;
    const float* caller;
    ~SupportSwap() { s->support = caller; }
};
}  // namespace

extern "C" int cmtip_acsolver_cg(cmtip_acsolver s, const float* d, float* x, double alpha, double rho,
                                 double phi, int maxiter, double rtol, int* iters, int* info, void* stream) {
    if (!valid(s)) return CMTIP_ERR_ARG;
    if (!d || !x || maxiter < 0) {
        set_error("acsolver_cg: bad arguments");
        return CMTIP_ERR_ARG;
    }
    hipStream_t caller = as_stream(stream);
    hipStream_t st = s->cg_stream;
    CMTIP_HIP(hipEventRecord(s->ev_in, caller));
    CMTIP_HIP(hipStreamWaitEvent(st, s->ev_in, 0));
    const int64_t n = (int64_t)s->M * s->M * s->M;
    double* sc = s->scal.as<double>();
    int* ctrl = s->ctrl.as<int>();
    const unsigned g = nblk(n, 2048);

    // the iterate and the support in solver-owned buffers: the captured block of
    // iterations refers to solver memory only, so it is kept across calls with the same
    // rho and phi (and re-captured when they change)
    CMTIP_TRY(s->xbuf.ensure(sizeof(float) * (size_t)n));
    float* xw = s->xbuf.as<float>();
    CMTIP_HIP(hipMemcpyAsync(xw, x, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, st));
    SupportSwap swap{s, s->support};
    if (swap.caller) {
        CMTIP_TRY(s->supbuf.ensure(sizeof(float) * (size_t)n));
        CMTIP_HIP(hipMemcpyAsync(s->supbuf.p, swap.caller, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, st));
        s->support = s->supbuf.as<const float>();
    }

    // b = d; ||b||, atol, early exit for b == 0; r = b - A x (A 0 = 0 exactly, so this is
    // scipy's "b - A x if x.any() else b"); rho = r.r
    s->cur_skip = nullptr;
    CMTIP_TRY(set_scalars(sc, 0, S_COUNT, 0.0, st));
    CMTIP_TRY(dot(s, d, d, n, sc, S_B2, st, nullptr));
    hipLaunchKernelGGL(cg_init_kernel, dim3(1), dim3(64), 0, st, sc, ctrl, maxiter, rtol);
    hipLaunchKernelGGL(copy_if_zero_rhs_kernel, dim3(g), dim3(kT), 0, st, d, xw, n, (const double*)sc);
    CMTIP_LAUNCH_CHECK();
    s->cur_skip = ctrl;
    int rc = matvec(s, xw, s->q.as<float>(), alpha, rho, phi, st);
    if (rc == CMTIP_OK) {
        hipLaunchKernelGGL(residual_kernel, dim3(g), dim3(kT), 0, st, d, (const float*)s->q.as<float>(),
                           s->r.as<float>(), n, (const int*)ctrl);
        rc = dot(s, s->r.as<const float>(), s->r.as<const float>(), n, sc, S_RHO, st, ctrl);
    }
    if (rc == CMTIP_OK) {
        // the loop check of the first iteration (each iteration ends with the next one's)
        hipLaunchKernelGGL(cg_check_kernel, dim3(1), dim3(64), 0, st, (const double*)sc, ctrl);
        if (hipGetLastError() != hipSuccess) {
            set_error("acsolver_cg: launch failed");
            rc = CMTIP_ERR_LAUNCH;
        }
    }
    // iterations: replay a captured block of kCgGraphIters iterations until the device
    // stops; the block is captured once per (rho, phi, support present) and kept
    const bool with_support = s->support != nullptr;
    if (s->cg_exec && (s->cg_rho != rho || s->cg_phi != phi || s->cg_with_support != with_support)) {
        (void)hipGraphExecDestroy(s->cg_exec);
        s->cg_exec = nullptr;
    }
    if (rc == CMTIP_OK && use_graphs() && maxiter > 0 && !s->cg_exec) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            int crc = CMTIP_OK;
            for (int k = 0; k < kCgGraphIters && crc == CMTIP_OK; ++k) crc = cg_iteration(s, xw, rho, phi, st);
            hipLaunchKernelGGL(publish_kernel, dim3(1), dim3(64), 0, st, (const int*)ctrl, (const double*)sc,
                               s->host_state_dev);
            const hipError_t e = hipStreamEndCapture(st, &graph);
            if (crc != CMTIP_OK || e != hipSuccess ||
                hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
                exec = nullptr;
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();  // a failed capture falls back to direct launches below
            set_error("");
        }
        s->cg_exec = exec;
        s->cg_rho = rho;
        s->cg_phi = phi;
        s->cg_with_support = with_support;
    }
    hipGraphExec_t exec = use_graphs() ? s->cg_exec : nullptr;
    // rounds launched before the host looks at the state: with rtol 0 the device can only
    // stop at maxiter (or at an exactly zero residual, after which the remaining
    // iterations do nothing), so all of them go at once; otherwise one per check
    const int all_rounds = (maxiter + kCgGraphIters - 1) / kCgGraphIters;
    volatile int* hs = s->host_state;
    for (int done = (maxiter == 0) ? 1 : 0, rounds = 0; rc == CMTIP_OK;) {
        const int batch = (rtol == 0.0 && rounds == 0) ? std::max(1, all_rounds) : 1;
        for (int b = 0; b < batch && rc == CMTIP_OK; ++b, ++rounds) {
            if (exec) {
                if (hipGraphLaunch(exec, st) != hipSuccess) {
                    set_error("acsolver_cg: graph launch failed");
                    rc = CMTIP_ERR_HIP;
                }
            } else {
                for (int k = 0; k < kCgGraphIters && rc == CMTIP_OK; ++k) rc = cg_iteration(s, xw, rho, phi, st);
                hipLaunchKernelGGL(publish_kernel, dim3(1), dim3(64), 0, st, (const int*)ctrl, (const double*)sc,
                                   s->host_state_dev);
            }
        }
        if (rc != CMTIP_OK) break;
        if (hipStreamSynchronize(st) != hipSuccess) {
            set_error("acsolver_cg: stream synchronisation failed");
            rc = CMTIP_ERR_HIP;
            break;
        }
        if (done || hs[0] || rounds > maxiter / kCgGraphIters + 2) break;
    }
    s->cur_skip = nullptr;
    if (rc == CMTIP_OK && hipMemcpyAsync(x, xw, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
        set_error("acsolver_cg: copy of the iterate failed");
        rc = CMTIP_ERR_HIP;
    }
    if (hipEventRecord(s->ev_out, st) != hipSuccess || hipStreamWaitEvent(caller, s->ev_out, 0) != hipSuccess) {
        set_error("acsolver_cg: stream ordering failed");
        if (rc == CMTIP_OK) rc = CMTIP_ERR_HIP;
    }
    if (rc != CMTIP_OK) return rc;
    if (maxiter == 0) {
        hipLaunchKernelGGL(publish_kernel, dim3(1), dim3(64), 0, st, (const int*)ctrl, (const double*)sc,
                           s->host_state_dev);
        CMTIP_HIP(hipStreamSynchronize(st));
    }
    if (iters) *iters = hs[2];
    if (info) *info = hs[1] ? 0 : maxiter;
    return CMTIP_OK;
}
