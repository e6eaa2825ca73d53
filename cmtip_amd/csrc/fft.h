// Mixed-radix Stockham FFT passes over one axis of a 3-D array, staged in LDS.
//
// A pass loads whole lines of the transform axis into LDS (several lines per
// workgroup; when the axis is strided the workgroup's lines are neighbours along
// the contiguous axis, so every global access is a contiguous run), runs all
// radix stages in LDS, and stores selected positions back.  Zero padding,
// cropping, fftshift/ifftshift, the NUFFT deconvolution weights, support masks,
// a mid-transform pointwise product (forward FFT -> multiply -> inverse FFT in
// one pass, used by the Toeplitz normal operator) and axpy-style epilogues are
// all folded into the load and store phases, so no pass needs a memset or a
// separate elementwise kernel.
#pragma once
#include "common.h"

namespace cmtip {

constexpr int kFftMaxStages = 20;

struct Fft1D {
    int n = 0;
    int nst = 0;
    int R[kFftMaxStages];
    int Ns[kFftMaxStages];
    int twoff[kFftMaxStages];  // twiddle table offset of each stage
    int omoff[kFftMaxStages];  // R-th roots of unity table offset (generic radices only)
    DevBuf tw;                 // float2 twiddles for the e^{-i} direction
    int ntw = 0;               // entries of tw
};

int fft1d_init(Fft1D& p, int n);
void fft1d_free(Fft1D& p);

// One FFT pass along one axis of a 3-D array.
//
// Lines are enumerated as (u, v), u in [0, nu), v in [0, nv); v should be the
// contiguous axis of the array when the transform axis is strided.  Element
// offsets (in elements, not bytes) are u*s_u + v*s_v + g*s_a.
//
// Input along the transform axis: LDS position p in [0, ia) reads global index
// ig1 + p; p in [n - ib, n) reads global ig2 + (p - (n - ib)); other positions
// are zero.  Output: p in [0, oa) is written to og1 + p and p in [n - ob, n) to
// og2 + (p - (n - ob)); other positions are dropped.
struct AxisDesc {
    const void* in = nullptr;
    int in_real = 0;  // 1: input is float (imaginary part zero), 0: float2
    int64_t in_sa = 0, in_su = 0, in_sv = 0;
    int ia = 0, ib = 0, ig1 = 0, ig2 = 0;
    const float* in_w = nullptr;     // per-LDS-position input weight (length n)
    const float* in_mask = nullptr;  // per-element input weight (float, input offsets)

    void* out = nullptr;
    int out_mode = 0;  // 0: float2, 1: real part as float, 2: imaginary part as float, 3: |value| as float
    int64_t out_sa = 0, out_su = 0, out_sv = 0;
    int oa = 0, ob = 0, og1 = 0, og2 = 0;
    const float* out_w = nullptr;     // per-LDS-position output weight (length n)
    float out_scale = 1.0f;
    const float* out_mask = nullptr;  // per-element output weight (float, output offsets)
    const float* add_src = nullptr;   // real mode only: out += add_scale * add_src[off]
    float add_scale = 0.0f;
    // Real mode only: a phase-retrieval update of the output array in place, with r the
    // real value after weights: 1 (error reduction) out = (S > 0 and r > 0) ? r : 0;
    // 2 (hybrid input-output) out = (S > 0 and r > 0) ? r : out - beta r; then where
    // r > rho_max, ER sets out = rho_max and HIO adds 2 beta r - rho_max.  S is
    // upd_support[off] (float 0 / 1).
    int upd_mode = 0;
    const float* upd_support = nullptr;
    float upd_beta = 0.0f;
    float upd_rho_max = 0.0f;

    int nu = 1, nv = 1;
    int sign = -1;  // -1: exp(-i...), +1: exp(+i...); both unnormalised

    // Optional fused pointwise product between a forward (sign) transform and an
    // inverse (-sign) transform of the same line: value *= mid[u*s_u+v*s_v+p*s_a].
    const float2* mid_c = nullptr;
    const float* mid_r = nullptr;
    // Other mid-pass maps (instead of the product): mid_amp replaces the magnitude,
    // value -> a value / |value| with a = mid_amp[off] (a zero value becomes (a, 0)),
    // and keeps the value where a < 0; mid_abs2 maps value -> (|value|^2, 0).
    const float* mid_amp = nullptr;
    int mid_abs2 = 0;
    int64_t mid_sa = 0, mid_su = 0, mid_sv = 0;

    // Hermitian half spectrum: with in_herm set, positions p in [n - ib, n) read
    // conj(in[n - p]) (ig2 unused), so a line stored as its first n/2 + 1 values is
    // extended to the full spectrum of a real signal.
    int in_herm = 0;

    // Load-phase CG direction update (real input only): p = r + beta p with
    // beta = cg_scal[0] / cg_scal[1], or p = r when cg_ctrl[2] == 1 (first iteration);
    // in points at r, the new p is written back to cg_p and transformed.
    float* cg_p = nullptr;
    const double* cg_scal = nullptr;
    const int* cg_ctrl = nullptr;

    // Real output modes: per-workgroup partial sums (double) of dot_src[off] * out[off]
    // over the values stored, written to dot_part[blockIdx.x]; fft_axis_grid gives the
    // number of workgroups.
    const float* dot_src = nullptr;
    double* dot_part = nullptr;

    // Optional device flag: when *skip != 0 the pass does nothing (lets a captured
    // sequence of passes be switched off from the device, e.g. after CG converged).
    const int* skip = nullptr;
};

int fft_axis_pass(const Fft1D& plan, const AxisDesc& d, hipStream_t stream);

// Number of workgroups fft_axis_pass launches for d (the length of dot_part it fills).
int64_t fft_axis_grid(const Fft1D& plan, const AxisDesc& d);

}  // namespace cmtip
