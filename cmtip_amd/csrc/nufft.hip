// 3-D type-1 / type-2 NUFFT plans, rotated-pixel coordinates and fused slicing.
//
// Type 2 (forward, cmtip/nufft/forward.py:13-82):
//   deconvolve + zero-pad + FFT (three LDS-staged axis passes, pruned to the
//   nonzero mode block) -> fine grid g (n^3) -> ES interpolation at every point.
// Type 1 (adjoint, cmtip/nufft/adjoint.py:13-86):
//   ES spreading of every point onto g -> FFT (three passes, pruned to the kept
//   modes) -> crop + deconvolve + support + scale folded into the last pass.
//
// Only the real part of a result is ever needed by the autocorrelation path
// (callers keep .real, cmtip/alignment/align.py:61-63, adjoint.py:40-43).  Because
// the interpolation weights are real, Re(interp(g)) = interp(Re g): the real-output
// type 2 keeps a real fine grid (half the gather bytes), and a real-strength type 1
// spreads into a real fine grid (half the atomics).
#include <math.h>

#include <algorithm>
#include <vector>

#include "es.h"
#include "fft.h"

using namespace cmtip;

struct cmtip_nufft_plan_s {
    int type = 0, M = 0, iflag = -1, w = 0, n = 0;
    double beta = 0.0, eps = 0.0;
    float tc_hi = 0.f, tc_lo = 0.f;  // n / (2 pi) split in two floats
    float beta_l2e = 0.f;            // beta * log2(e), the kernels' exponent scale
    EsKernel esk{};                  // beta_l2e and the tap polynomials, passed to the kernels
    Fft1D fft;
    DevBuf deconv;   // n floats
    DevBuf fine;     // n^3 float2 (used as n^3 float when real)
    DevBuf fine64;   // n^3 int64: fixed-point fine grid of real spreads (converted into fine)
    DevBuf tmpA;     // M * n * n float2
    DevBuf tmpB;     // M * M * n float2
    DevBuf rinv;     // 9 doubles per orientation
    DevBuf part;     // slice-moment partials: 2 doubles per tile (or per wave), then the reduction stage
    int slice_ready = 0;
    int64_t npts = 0;
    const float* x = nullptr;
    const float* y = nullptr;
    const float* z = nullptr;
    // spatial bins of the points (set by setpts): counting sort into bins of bs^3 fine
    // cells, bin-sorted coordinate copies, and the chunk list worked by the kernels
    int bs = 0, nbx = 0, chunk = 0;
    int64_t nbins = 0;
    // requested launch settings (0 = default): bin edge, points per chunk, persistent
    // workgroups per CU of the spread / interpolation kernels
    int bs_req = 0, chunk_req = 0, wpc_req = 0;
    // slicing route: requested (0 auto, 1 fused tiles, 2 binned points, 3 binned tiles)
    // and last used; the automatic choice's probe (shares of tiles whose fused box fits
    // LDS and whose footprint fits the tile route) is cached per detector geometry
    int route_req = 0, route_used = 0, route_probed = 0;
    double fit_box = 0.0, fit_tile = 0.0;
    const double* route_s = nullptr;
    const int* route_order = nullptr;
    int64_t route_P = -1;
    float route_ext = 0.f;
    DevBuf probe;  // 3 int32: tiles staged in LDS, tiles probed, tiles within the tile margin
    // kernel timer (cmtip_slice_timer): event pairs around the main slicing kernel
    int timing = 0, timed = 0;
    hipEvent_t tev[2 * 64] = {};
    DevBuf amax;      // 1 uint32: bits of max |coordinate| of the registered points (setpts)
    int amax_valid = 0;  // 1: amax describes the points with records; 2: none registered (0)
    DevBuf bin_of;    // N int32
    DevBuf count;     // nbins int32 (also the scatter cursors)
    DevBuf offset;    // nbins + 1 int32
    DevBuf nchunk;    // nbins + 1 int32: chunk counts, then their exclusive scan
    DevBuf perm;      // N int32: sorted position -> original index
    DevBuf sorted;    // N float4 (x, y, z, original index as int bits) in sorted order
    DevBuf chunks;    // max_chunks x int4 (bin, start, end, 0)
    DevBuf ctrl;      // int32: [0] number of chunks, [1] work counter, [2] bits of max |c|, [3] some c < 0
    int64_t max_chunks = 0;
    int bound = 0;    // 1: the records carry real strengths (setpts_strengths), not indices
    int rot_records = 0;  // 1: the records hold rotated pixels (setpts_rotated), no arrays registered
};

namespace {

constexpr int kBlock = 256;
constexpr int kMomentBlocks = 256;  // first-stage blocks of the fixed-order moment reduction

// ---------------------------------------------------------------- device helpers

// phi(z) = exp(beta (sqrt(1 - (2z/w)^2) - 1)) with the hardware square root and
// base-2 exponential (v_sqrt_f32, v_exp_f32; about 1 ulp each).  beta_l2e = beta log2(e).
__device__ __forceinline__ float es_phi(float z, float two_over_w, float beta_l2e) {
    const float u = z * two_over_w;
    const float s = fmaf(-u, u, 1.0f);
    const float r = __builtin_amdgcn_sqrtf(fmaxf(s, 0.0f));
    const float e = __builtin_amdgcn_exp2f(beta_l2e * (r - 1.0f));
    return s > 0.0f ? e : 0.0f;
}

// Tap weights of one axis in tap order, wt[a] = phi(a - tau) with tau = t - l0 in
// (W/2 - 1, W/2]: for W <= kEsPolyMaxW from the tap polynomials (even part E(y^2) and
// odd part y O(y^2) per mirror pair, y = tau - (W - 1)/2; full-rate FMAs only), else
// from es_phi.
template <int W>
__device__ __forceinline__ void es_taps(float tau, const EsKernel& K, float* wt) {
    if constexpr (W <= kEsPolyMaxW) {
        constexpr int NE = (W + 1) / 2 + 1, NO = (W + 2) / 2;
        const float y = tau - 0.5f * (float)(W - 1);
        const float y2 = y * y;
#pragma unroll
        for (int j = 0; j < W / 2; ++j) {
            const float* c = K.c + j * (NE + NO);
            float e = c[NE - 1];
#pragma unroll
            for (int k = NE - 2; k >= 0; --k) e = fmaf(e, y2, c[k]);
            float o = c[NE + NO - 1];
#pragma unroll
            for (int k = NO - 2; k >= 0; --k) o = fmaf(o, y2, c[NE + k]);
            o *= y;
            wt[j] = e + o;
            wt[W - 1 - j] = e - o;
        }
        if constexpr ((W & 1) != 0) {
            const float* c = K.c + (W / 2) * (NE + NO);
            float e = c[NE - 1];
#pragma unroll
            for (int k = NE - 2; k >= 0; --k) e = fmaf(e, y2, c[k]);
            wt[W / 2] = e;
        }
    } else {
        const float tw = 2.0f / (float)W;
#pragma unroll
        for (int a = 0; a < W; ++a) wt[a] = es_phi((float)a - tau, tw, K.beta_l2e);
    }
}

// Fine-grid coordinate t = x n / (2 pi) with the constant split in two floats, so the
// product carries about one rounding.  Non-finite or huge inputs are clamped so every
// index below stays in range whatever the caller passes.
// l mod n (result in [0, n)) for |l| < 2^22, n >= 1: the floor of l / n from a float
// estimate, which is within 1/2 of the true quotient (relative error <= 2^-23), then
// one correction each way; about half the instructions of an integer % by a runtime n.
__device__ __forceinline__ int mod_small(int l, int n) {
    const int q = (int)floorf((float)l * (1.0f / (float)n));
    int r = l - q * n;
    r += r < 0 ? n : 0;
    r -= r >= n ? n : 0;
    return r;
}

// l / d for 0 <= l < 2^22, d >= 1 (same estimate and corrections)
__device__ __forceinline__ int div_small(int l, int d) {
    int q = (int)((float)l * (1.0f / (float)d));
    const int r = l - q * d;
    q += r >= d ? 1 : 0;
    q -= r < 0 ? 1 : 0;
    return q;
}

__device__ __forceinline__ float fine_coord(float x, float hi, float lo, int n) {
    float t = fmaf(x, hi, x * lo);
    const float lim = 4.0f * (float)n;
    t = fminf(fmaxf(t, -lim), lim);
    return t;
}

// ---------------------------------------------------------------- binning
//
// setpts sorts the points by the fine-grid bin holding the first node of their kernel
// footprint (l0 mod n, per axis): bins are bs^3 cells, so every footprint of a bin's
// points lies in its subgrid of (bs + W - 1)^3 cells.  Bins are cut into chunks of at
// most chunk <= kChunk points; each chunk is one unit of work for the spread (type 1) and
// interpolation (type 2) kernels, which stage the chunk's subgrid in LDS.

constexpr int kChunk = 4096;  // largest chunk (the fixed-point spreading margin assumes it)
constexpr int kSpreadBlock = 256;  // interpolation workgroups
// Spreading workgroups of 512 threads (three per CU with their LDS) and points dealt
// to lanes by bank class: measured over 81.92 M points, type 1 at M = 81 13.7 ms
// (256 threads, no dealing) -> 11.2 (256, dealt) -> 9.25 ms (512, dealt); unit
// strengths at 162 modes 9.8 -> 9.1 ms.
#ifndef CMTIP_SPREAD_THREADS
#define CMTIP_SPREAD_THREADS 512
#endif
constexpr int kSpreadThreads = CMTIP_SPREAD_THREADS;  // spreading workgroups
#ifndef CMTIP_SPREAD_DEAL
#define CMTIP_SPREAD_DEAL 1
#endif
constexpr bool kSpreadDeal = CMTIP_SPREAD_DEAL;  // deal points to lanes by LDS bank class
// a 64-bit accumulator sits on one of the 16 pairs of the 32 four-byte LDS banks;
// dealing by those 16 classes measured 8.4-8.5 ms at M = 81 (81.92 M points, real
// strengths) against 9.0 (32 classes), 10.0 (8) and 11.0 (64)
constexpr int kSpreadClasses = 16;
// 32-bit accumulators (unit strengths): one class per bank, 32 classes, so that the 32
// lanes of each pass of a 32-bit LDS atomic hit 32 distinct banks
constexpr int kSpreadClasses32 = 32;
// static LDS of the spreading kernel (point dealing), beside its dynamic subgrid
constexpr size_t kSpreadStaticLds = 3 * kChunk + 4 * 3 * kSpreadClasses32 + 64;
constexpr size_t kMaxLds = 160 * 1024;     // LDS per workgroup on gfx950
constexpr size_t kDefaultLds = 64 * 1024;  // dynamic LDS allowed without raising the limit
constexpr size_t kStaticLds = 256;         // reserve for the binned kernels' static LDS

// Default subgrids are 16^3 cells for w <= 12 and 20^3 above (bin edge = subgrid edge -
// w + 1); a plan can be given another bin edge, chunk size and launch width
// (cmtip_nufft_set_bins, chosen by cmtip_nufft_autotune).
__host__ __device__ inline int bin_size_for(int w) { return w <= 12 ? 17 - w : 21 - w; }

// Row reads of the interpolation kernels (WideRow): 4 = quad + pair + single when
// W = 7 (else as 2), 2 = pairs + singles, 1 = single floats.  Measured on gfx950 at
// M = 81: binned interpolation 5.06 ms (1) -> 3.42 (2) -> 2.69 (4) for 81.92 M points;
// the fused tile kernel, bound by staging its tile boxes, 10.9 (1) / 10.5 (2) / 11.3 (4)
// ms for 163.8 M points.
#ifndef CMTIP_INTERP_READS
#define CMTIP_INTERP_READS 4
#endif
#ifndef CMTIP_SLICE_READS
#define CMTIP_SLICE_READS 2
#endif

// Row pitch SP and plane pitch SSP of the interpolation subgrid (S^3 real cells,
// layout x * SSP + y * SP + z).  Subgrids are read in wide rows: both pitches are
// multiples of the row alignment (4 for w = 7, else 2), padded by one alignment unit so
// that consecutive rows and planes start in different LDS banks.
__host__ __device__ inline void interp_pitch(int S, int W, int* SP, int* SSP) {
    const int A = CMTIP_INTERP_READS == 1 ? 1 : ((W == 7 && CMTIP_INTERP_READS == 4) ? 4 : 2);
    *SP = (S + A - 1) / A * A + A;
    *SSP = S * *SP + A;
}

// Row / plane pitch (in cells) of the spreading subgrid of 64-bit accumulators: odd
// row pitch, plane pitch padded by 5 cells so that neighbouring planes start in other
// banks (measured over 81.92 M points: type 1 at M = 81 17.2 ms with no plane pad,
// 15.9 / 14.1 / 13.9 ms with pads 1 / 3 / 5; at 162 modes 11.6 -> 9.8 ms).
#ifndef CMTIP_SPREAD_SP_PAD
#define CMTIP_SPREAD_SP_PAD 1
#endif
#ifndef CMTIP_SPREAD_SSP_PAD
#define CMTIP_SPREAD_SSP_PAD 5
#endif
__host__ __device__ inline void spread_pitch(int S, int* SP, int* SSP) {
    *SP = S + CMTIP_SPREAD_SP_PAD;
    *SSP = S * *SP + CMTIP_SPREAD_SSP_PAD;
}

// LDS bytes of the spread (64-bit accumulators, two per cell when complex) or the
// interpolation (float) subgrid.
inline size_t subgrid_lds(int bs, int W, bool spread, bool cplx) {
    const int S = bs + W - 1;
    int SP, SSP;
    if (spread) {
        spread_pitch(S, &SP, &SSP);
        return (size_t)S * SSP * 8 * (cplx ? 2 : 1);
    }
    interp_pitch(S, W, &SP, &SSP);
    return (size_t)S * SSP * 4;
}

__device__ __forceinline__ int first_node_wrapped(float x, float hi, float lo, int n, int W, float* t_out) {
    const float t = fine_coord(x, hi, lo, n);
    *t_out = t;
    // |t| <= 4n, so |l0| < 2^22 for every grid a plan allows
    const int l0 = (int)ceilf(t - 0.5f * (float)W);
    return mod_small(l0, n);
}

// Wave-aggregated counter update: lanes of a wave holding the same bin issue one
// atomicAdd for the group and get consecutive slots (points that are neighbours in
// the input, e.g. adjacent pixels, mostly share a bin).  Returns the lane's slot.
__device__ __forceinline__ int wave_bin_add(int* __restrict__ counter, int b, bool valid) {
    const int lane = threadIdx.x & 63;
    unsigned long long active = __ballot(valid);
    int slot = 0;
    while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const int lb = __shfl(b, leader, 64);
        const unsigned long long m = __ballot(valid && b == lb) & active;
        int old = 0;
        if (lane == leader) old = atomicAdd(counter + lb, __popcll(m));
        old = __shfl(old, leader, 64);
        if ((m >> lane) & 1ull) slot = old + __popcll(m & ((1ull << lane) - 1ull));
        active &= ~m;
    }
    return slot;
}

// Point sources of the binning and interpolation kernels.  PtsArrays: registered
// coordinate arrays.  PtsRotated (defined with the rotation helpers below): the
// rotated pixels of the slicing route, point j = o * P + slot at pixel order[slot],
// generated from (rinv, s) wherever they are needed instead of being stored.
struct PtsArrays {
    const float* x;
    const float* y;
    const float* z;
    __device__ __forceinline__ void at(int64_t j, float* px, float* py, float* pz) const {
        *px = x[j];
        *py = y[j];
        *pz = z[j];
    }
};

// |x| for the coordinate range check, NaN counted as out of range (+inf)
__device__ __forceinline__ float range_abs(float v) { return v == v ? fabsf(v) : __int_as_float(0x7f800000); }


// Workgroup maxima of non-negative float bits (and an OR of flags) through LDS, then one
// global atomic per workgroup: per-wave atomics on one address serialise (16 k of them
// cost ~0.3 ms per pass over 81.92 M points).
__device__ __forceinline__ void block_max_or(float m, unsigned flag, unsigned* gmax, unsigned* gflag) {
    __shared__ unsigned s_max, s_flag;
    if (threadIdx.x == 0) {
        s_max = 0u;
        s_flag = 0u;
    }
    __syncthreads();
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    const bool any = __ballot(flag != 0u) != 0;
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&s_max, __float_as_uint(m));
        if (any) atomicOr(&s_flag, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (gmax) atomicMax(gmax, s_max);
        if (gflag && s_flag) atomicOr(gflag, 1u);
    }
}

template <class Src>
__global__ void bin_count_kernel(Src pts, int64_t N, int n, int W, int bs, int nbx, float hi, float lo,
                                 int* __restrict__ bin_of, int* __restrict__ count, unsigned* __restrict__ amax,
                                 const float* __restrict__ cstr, unsigned* __restrict__ cmax) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool valid = j < N;
    int b = 0;
    float m = 0.0f, cm = 0.0f;
    bool neg = false;
    if (valid) {
        float t, x, y, z;
        pts.at(j, &x, &y, &z);
        if (amax) m = fmaxf(range_abs(x), fmaxf(range_abs(y), range_abs(z)));
        if (cstr) {
            cm = fabsf(cstr[j]);
            neg = cstr[j] < 0.0f;
        }
        const int bx = div_small(first_node_wrapped(x, hi, lo, n, W, &t), bs);
        const int by = div_small(first_node_wrapped(y, hi, lo, n, W, &t), bs);
        const int bz = div_small(first_node_wrapped(z, hi, lo, n, W, &t), bs);
        b = (bx * nbx + by) * nbx + bz;
        bin_of[j] = b;
    }
    if (amax) block_max_or(m, 0u, amax, nullptr);
    if (cstr) block_max_or(cm, neg ? 1u : 0u, cmax, cmax + 1);
    wave_bin_add(count, b, valid);
}

// One workgroup: offset = exclusive scan of count, nchunk = exclusive scan of
// ceil(count / chunk); ctrl[0] = total chunks.  Counts are then cleared for use as
// scatter cursors.  Each thread sums a contiguous segment of the bins, one shuffle scan
// over the 1024 segment sums gives each segment its start, and the thread writes its
// segment's prefixes (integer sums: the result does not depend on the split).
__global__ void __launch_bounds__(1024) bin_scan_kernel(int* __restrict__ count, int64_t nbins, int chunk,
                                                        int* __restrict__ offset, int* __restrict__ nchunk,
                                                        int* __restrict__ ctrl) {
    __shared__ int wa[16], wb[16];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t per = (nbins + 1023) / 1024;
    const int64_t b0 = min(nbins, (int64_t)t * per), b1 = min(nbins, b0 + per);
    int sa = 0, sb = 0;
#pragma unroll 8
    for (int64_t i = b0; i < b1; ++i) {
        const int c = count[i];
        sa += c;
        sb += (c + chunk - 1) / chunk;
    }
    int ia = sa, ib = sb;  // inclusive scan over the wave
    for (int off = 1; off < 64; off <<= 1) {
        const int ta = __shfl_up(ia, off, 64), tb = __shfl_up(ib, off, 64);
        if (lane >= off) {
            ia += ta;
            ib += tb;
        }
    }
    if (lane == 63) {
        wa[wave] = ia;
        wb[wave] = ib;
    }
    __syncthreads();
    if (wave == 0) {  // inclusive scan of the 16 wave totals
        int xa = lane < 16 ? wa[lane] : 0, xb = lane < 16 ? wb[lane] : 0;
        for (int off = 1; off < 16; off <<= 1) {
            const int ta = __shfl_up(xa, off, 64), tb = __shfl_up(xb, off, 64);
            if (lane >= off) {
                xa += ta;
                xb += tb;
            }
        }
        if (lane < 16) {
            wa[lane] = xa;
            wb[lane] = xb;
        }
    }
    __syncthreads();
    int ea = ia - sa + (wave ? wa[wave - 1] : 0), eb = ib - sb + (wave ? wb[wave - 1] : 0);
#pragma unroll 8
    for (int64_t i = b0; i < b1; ++i) {
        const int c = count[i];
        offset[i] = ea;
        nchunk[i] = eb;
        count[i] = 0;
        ea += c;
        eb += (c + chunk - 1) / chunk;
    }
    if (t == 1023) {  // its running prefix ends at the total whether or not its segment is empty
        offset[nbins] = ea;
        nchunk[nbins] = eb;
        ctrl[0] = eb;
    }
}

// COPY: write one 16-byte record (x, y, z, original index) per sorted position, one
// scattered store per point, the coordinates from the point source (registered arrays,
// or rotated pixels generated again); else (the slicing route, which regenerates the
// coordinates when it interpolates) write the original index to perm.
template <bool COPY, class Src>
__global__ void bin_scatter_kernel(Src pts, int64_t N, const int* __restrict__ bin_of,
                                   const int* __restrict__ offset, int* __restrict__ cursor, int* __restrict__ perm,
                                   float4* __restrict__ sorted, const float* __restrict__ cstr) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool valid = j < N;
    const int b = valid ? bin_of[j] : 0;
    const int slot = wave_bin_add(cursor, b, valid);
    if (!valid) return;
    const int pos = offset[b] + slot;
    if (COPY) {
        float x, y, z;
        pts.at(j, &x, &y, &z);
        sorted[pos] = make_float4(x, y, z, cstr ? cstr[j] : __int_as_float((int)j));
    } else {
        perm[pos] = (int)j;
    }
}

// Workgroup-privatised binning (when the bin counters fit LDS): a workgroup takes a
// contiguous range of points, counts them in an LDS histogram and adds only its
// non-zero bins to the global counters; the scatter pass counts again, reserves one
// contiguous range per (workgroup, bin) with a single global atomic and hands out the
// slots from LDS.  Global atomics drop from one per (wave, bin) to one per
// (workgroup, bin).
constexpr int kBinBlock = 1024;
constexpr int64_t kLocalBinsMax = 36864;  // 144 KB of LDS counters

// cstr (optional): real strengths bound to the points at setpts; their max |c| and
// whether any is negative go to cmax[0], cmax[1] (the spread's fixed-point scale)
template <class Src>
__global__ void __launch_bounds__(kBinBlock) bin_count_local_kernel(Src pts, int64_t N, int64_t per_block, int n,
                                                                   int W, int bs, int nbx, int nbins, float hi,
                                                                   float lo, int* __restrict__ bin_of,
                                                                   int* __restrict__ count,
                                                                   unsigned* __restrict__ amax,
                                                                   const float* __restrict__ cstr,
                                                                   unsigned* __restrict__ cmax) {
    extern __shared__ int hist[];
    for (int i = threadIdx.x; i < nbins; i += kBinBlock) hist[i] = 0;
    __syncthreads();
    const int64_t j0 = (int64_t)blockIdx.x * per_block, j1 = min(N, j0 + per_block);
    float m = 0.0f, cm = 0.0f;
    bool neg = false;
    for (int64_t j = j0 + threadIdx.x; j < j1; j += kBinBlock) {
        float t, x, y, z;
        pts.at(j, &x, &y, &z);
        if (amax) m = fmaxf(m, fmaxf(range_abs(x), fmaxf(range_abs(y), range_abs(z))));
        if (cstr) {
            const float v = cstr[j];
            cm = fmaxf(cm, fabsf(v));
            neg = neg || v < 0.0f;
        }
        const int bx = div_small(first_node_wrapped(x, hi, lo, n, W, &t), bs);
        const int by = div_small(first_node_wrapped(y, hi, lo, n, W, &t), bs);
        const int bz = div_small(first_node_wrapped(z, hi, lo, n, W, &t), bs);
        const int b = (bx * nbx + by) * nbx + bz;
        bin_of[j] = b;
        atomicAdd(hist + b, 1);
    }
    if (amax) block_max_or(m, 0u, amax, nullptr);
    if (cstr) block_max_or(cm, neg ? 1u : 0u, cmax, cmax + 1);
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += kBinBlock)
        if (hist[i]) atomicAdd(count + i, hist[i]);
}

// COPY: one 16-byte record per point, (x, y, z, original index as int bits), or with
// cstr (x, y, z, strength); the coordinates from the point source
template <bool COPY, class Src>
__global__ void __launch_bounds__(kBinBlock) bin_scatter_local_kernel(
    Src pts, int64_t N, int64_t per_block, int nbins, const int* __restrict__ bin_of, const int* __restrict__ offset,
    int* __restrict__ cursor, int* __restrict__ perm, float4* __restrict__ sorted, const float* __restrict__ cstr) {
    extern __shared__ int hist[];
    for (int i = threadIdx.x; i < nbins; i += kBinBlock) hist[i] = 0;
    __syncthreads();
    const int64_t j0 = (int64_t)blockIdx.x * per_block, j1 = min(N, j0 + per_block);
    for (int64_t j = j0 + threadIdx.x; j < j1; j += kBinBlock) atomicAdd(hist + bin_of[j], 1);
    __syncthreads();
    // the workgroup's range of each bin: hist[b] becomes its first position
    for (int i = threadIdx.x; i < nbins; i += kBinBlock) {
        const int c = hist[i];
        if (c) hist[i] = offset[i] + atomicAdd(cursor + i, c);
    }
    __syncthreads();
    for (int64_t j = j0 + threadIdx.x; j < j1; j += kBinBlock) {
        const int pos = atomicAdd(hist + bin_of[j], 1);
        if (COPY) {
            float x, y, z;
            pts.at(j, &x, &y, &z);
            sorted[pos] = make_float4(x, y, z, cstr ? cstr[j] : __int_as_float((int)j));
        } else {
            perm[pos] = (int)j;
        }
    }
}

__global__ void chunk_fill_kernel(const int* __restrict__ offset, const int* __restrict__ nchunk, int64_t nbins,
                                  int chunk, int4* __restrict__ chunks) {
    const int64_t b = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (b >= nbins) return;
    const int start = offset[b], end = offset[b + 1];
    int c = nchunk[b];
    for (int s0 = start; s0 < end; s0 += chunk, ++c) chunks[c] = make_int4((int)b, s0, min(end, s0 + chunk), 0);
}

// ---------------------------------------------------------------- binned spreading / interpolation
//
// Persistent workgroups take chunks from a device work counter until none are left.
// Spreading: zero the LDS subgrid, add every point's w^3 contributions with LDS
// atomics, then add the nonzero subgrid cells to the fine grid with one global atomic
// each.  Interpolation: load the subgrid from the fine grid, then every point gathers
// its w^3 taps from LDS and is written to its original position.

template <bool CPLX>
struct Cell;
template <>
struct Cell<false> {
    typedef float T;
};
template <>
struct Cell<true> {
    typedef float2 T;
};

template <int W>
__device__ __forceinline__ void local_weights(float t, int l0w, int origin, const EsKernel& K, int* loc,
                                              float* wt) {
    // l0w: wrapped first node; t: fine coordinate (unwrapped); origin: bin origin cell
    const float l0 = ceilf(t - 0.5f * (float)W);
    *loc = l0w - origin;
    es_taps<W>(t - l0, K, wt);
}

// ---------------------------------------------------------------- wide LDS row reads
//
// Interpolation reads W consecutive floats of an LDS row for every (x, y) of a point's
// footprint.  With the row and plane pitches multiples of kAlign, all row starts of a
// lane share one alignment m = start mod kAlign, and every row is read with aligned
// wide loads: for W = 7 one b128, one b64 and one b32 read whatever m is; for other
// widths b64 pairs and one or two b32 reads.  The lane evaluates its z weights in the
// order the values arrive (load_offsets), so rows need no shuffling.  With scattered
// row starts, wide reads move about twice as many bytes per LDS cycle as single floats
// (measured on gfx950: 36 B/clk per CU for b32, 62 for b64, 76 for b128).
template <int W, int READS = 4>
struct WideRow {
    static constexpr bool kQuad = (W == 7) && READS == 4;
    static constexpr bool kSingles = READS == 1;
    static constexpr int kAlign = kQuad ? 4 : (kSingles ? 1 : 2);
    static constexpr int kPairs = kQuad ? 1 : (kSingles ? 0 : ((W & 1) ? (W - 1) / 2 : W / 2 - 1));
    int oq, op, os;  // offsets from the row start of the quad (W = 7), first pair, single
    __device__ __forceinline__ explicit WideRow(int start) {
        const int m = start & (kAlign - 1);
        if (kSingles) {
            oq = op = os = 0;
        } else if (kQuad) {
            // m = 0: [0..3] [4,5] [6];  1: [3..6] [1,2] [0];  2: [2..5] [0,1] [6];  3: [1..4] [5,6] [0]
            oq = (4 - m) & 3;
            op = (0x5014 >> (4 * m)) & 15;
            os = (m & 1) ? 0 : 6;
        } else {
            oq = 0;
            op = m;
            os = m ? 0 : W - ((W & 1) ? 1 : 2);
        }
    }
    // z offsets (0 .. W-1) of the values in load order
    __device__ __forceinline__ void load_offsets(int* z) const {
        if (kSingles) {
#pragma unroll
            for (int i = 0; i < W; ++i) z[i] = i;
            return;
        }
        int k = 0;
        if (kQuad)
#pragma unroll
            for (int i = 0; i < 4; ++i) z[k++] = oq + i;
#pragma unroll
        for (int i = 0; i < kPairs; ++i) {
            z[k++] = op + 2 * i;
            z[k++] = op + 2 * i + 1;
        }
        z[k++] = os;
        if (!kQuad && !(W & 1)) z[k++] = W - 1;
    }
    // W = 7 quad rows: tap-order weights wt into load order (the quad is taps oq .. oq + 3,
    // oq in {0, 1, 2, 3}; the pair op, op + 1 with op in {0, 1, 4, 5}; the single 0 or 6)
    __device__ __forceinline__ void pick(const float* wt, float* wl) const {
        const bool q1 = (oq & 1) != 0, q2 = (oq & 2) != 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float s0 = q1 ? wt[i + 1] : wt[i];
            const float s1 = q1 ? wt[i + 3] : wt[i + 2];
            wl[i] = q2 ? s1 : s0;
        }
        const bool p1 = (op & 1) != 0, p4 = (op & 4) != 0;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const float s0 = p1 ? wt[i + 1] : wt[i];
            const float s1 = p1 ? wt[i + 5] : wt[i + 4];
            wl[4 + i] = p4 ? s1 : s0;
        }
        wl[6] = os ? wt[6] : wt[0];
    }
    // sum_k wl[k] * value_k over the row starting at row (16-byte aligned base + start)
    __device__ __forceinline__ float dot(const float* row, const float* wl) const {
        float q = 0.f;
        if (kSingles) {
#pragma unroll
            for (int i = 0; i < W; ++i) q = fmaf(wl[i], row[i], q);
            return q;
        }
        int k = 0;
        if (kQuad) {
            const float4 a = *reinterpret_cast<const float4*>(row + oq);
            q = wl[0] * a.x;
            q = fmaf(wl[1], a.y, q);
            q = fmaf(wl[2], a.z, q);
            q = fmaf(wl[3], a.w, q);
            k = 4;
        }
#pragma unroll
        for (int i = 0; i < kPairs; ++i) {
            const float2 b = *reinterpret_cast<const float2*>(row + op + 2 * i);
            q = fmaf(wl[k], b.x, q);
            q = fmaf(wl[k + 1], b.y, q);
            k += 2;
        }
        q = fmaf(wl[k], row[os], q);
        if (!kQuad && !(W & 1)) q = fmaf(wl[k + 1], row[W - 1], q);
        return q;
    }
};

// ES weights of one axis in the load order of a WideRow: the tap polynomials picked
// into that order (W = 7 quad rows), else es_phi at the load-order offsets.
template <int W, int READS>
__device__ __forceinline__ void weights_at(float t, const WideRow<W, READS>& wr, const EsKernel& K, float* wl) {
    const float tau = t - ceilf(t - 0.5f * (float)W);
    if constexpr (WideRow<W, READS>::kQuad && W <= kEsPolyMaxW) {
        float wt[W];
        es_taps<W>(tau, K, wt);
        wr.pick(wt, wl);
    } else {
        int z[W];
        wr.load_offsets(z);
        const float tw = 2.0f / (float)W;
#pragma unroll
        for (int k = 0; k < W; ++k) wl[k] = es_phi((float)z[k] - tau, tw, K.beta_l2e);
    }
}

// Spreading accumulates in LDS in 64-bit fixed point: integer LDS atomics are an
// order of magnitude faster than float ones on this hardware, and integer sums are
// exact and independent of the order of the adds.  Every contribution (an FP32
// product) is scaled by 2^sc and rounded to a 32-bit integer (one rounding and one
// conversion instruction, plus a sign extension for signed strengths); sc is chosen from
// max|c| so that every contribution lies below 2^30 in magnitude.  A chunk of at most
// kChunk = 2^12 points then sums to below 2^42 in any cell, and the 64-bit fine grid of a
// real spread holds 2^33 points' contributions (beyond that the caller lowers sc).
__device__ __forceinline__ int spread_scale_exp(float cmax) {
    if (!(cmax > 0.0f)) return 0;
    int e;
    frexpf(cmax, &e);  // cmax < 2^e
    return 30 - e;
}

// out_bits[0] = bits of max |c| (non-negative floats order like their bits; NaNs are
// ignored, as fmaxf does); out_bits[1] is set to 1 when some c < 0 (the caller clears
// both).  16-byte aligned input is read as float4, four loads in flight per thread.
__global__ void absmax_kernel(const float* __restrict__ c, int64_t count, unsigned* __restrict__ out_bits) {
    float m = 0.0f;
    bool neg = false;
    const int64_t tid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * kBlock;
    auto take = [&](float v) {
        m = fmaxf(m, fabsf(v));
        neg = neg || v < 0.0f;
    };
    int64_t done = 0;
    if ((reinterpret_cast<uintptr_t>(c) & 15) == 0) {
        const float4* c4 = reinterpret_cast<const float4*>(c);
        const int64_t n4 = count >> 2;
        int64_t i = tid;
        for (; i + 3 * nth < n4; i += 4 * nth) {
            float4 v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = c4[i + k * nth];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                take(v[k].x);
                take(v[k].y);
                take(v[k].z);
                take(v[k].w);
            }
        }
        for (; i < n4; i += nth) {
            const float4 v = c4[i];
            take(v.x);
            take(v.y);
            take(v.z);
            take(v.w);
        }
        done = n4 << 2;
    }
    for (int64_t i = done + tid; i < count; i += nth) take(c[i]);
    block_max_or(m, neg ? 1u : 0u, out_bits, out_bits + 1);
}

// Contribution w p (|w p| < 2^30, w >= 0) of a spread as a 64-bit fixed-point addend:
// the nearest integer (ties away from zero) as trunc(w p + h) with h = +-1/2 of the sign
// of p, one fused multiply-add and one 32-bit conversion, sign-extended (zero-extended
// for non-negative strengths, h = 1/2).  Above 2^23 the product is already an integer
// and adding h leaves it unchanged.
template <bool NONNEG>
__device__ __forceinline__ unsigned long long fixed_addend(float w, float p, float h) {
    if (NONNEG) return (unsigned long long)(uint32_t)fmaf(w, p, 0.5f);
    return (unsigned long long)(long long)(int)fmaf(w, p, h);
}

// U32 (unit strengths only): 32-bit LDS accumulators at the scale 2^kOnes32Scale per
// contribution (every weight product is <= 1, so a chunk's kChunk = 2^12 contributions to
// a cell stay below 2^31), added to the 64-bit fine grid shifted to its scale (which is
// at least 2^kOnes32Scale while npts < 2^43): exact, so
// the result is the same as the 64-bit path's up to the rounding of each weight product
// to 2^-19 (at most 1e-6 per tap), and deterministic.  Half the LDS per subgrid and
// twice the atomic rate of 64-bit LDS atomics.
constexpr int kOnes32Scale = 19;

template <int W, bool CPLX, bool U32 = false>
__global__ void __launch_bounds__(kSpreadThreads) spread_binned_kernel(
    typename Cell<CPLX>::T* __restrict__ grid, int n, const EsKernel esk, float hi, float lo, int bs, int nbx,
    const float4* __restrict__ sorted, const typename Cell<CPLX>::T* __restrict__ c,
    const unsigned* __restrict__ cmax_bits,
    const int4* __restrict__ chunks, int* __restrict__ ctrl, int bound, unsigned long long* __restrict__ fixed,
    int shift) {
    static_assert(!U32 || !CPLX, "32-bit accumulators for real unit strengths");
    constexpr int NC = CPLX ? 2 : 1;  // integer accumulators per cell
    using AccT = typename std::conditional<U32, unsigned, unsigned long long>::type;
    extern __shared__ unsigned long long acc_raw[];
    AccT* acc = reinterpret_cast<AccT*>(acc_raw);
    // LDS layout x * SSP + y * SP + z (spread_pitch)
    const int S = bs + W - 1, S3 = S * S * S;
    int SP, SSP;
    spread_pitch(S, &SP, &SSP);
    const int L3 = S * SSP;
    __shared__ int chunk_id, rounds;
    constexpr int NCL = U32 ? kSpreadClasses32 : kSpreadClasses;  // bank classes
    __shared__ int cls_count[NCL], cls_start[NCL], cls_fill[NCL];
    __shared__ uint16_t dealt[kChunk];  // chunk-local point indices grouped by class
    __shared__ uint8_t cls_of[kChunk];
    const int total = ctrl[0];
    // bound: real strengths carried in the records (setpts_strengths); their max |c| and
    // sign were recorded at setpts
    const bool has_c = c != nullptr || bound;
    // fixed (real spreads): the chunks' integer sums are added into a 64-bit fine grid,
    // exactly and in any order, so the scale leaves room for all points (shift)
    const int sc = spread_scale_exp(has_c ? __uint_as_float(*cmax_bits) : 1.0f) - (fixed ? shift : 0);
    const bool nonneg = !CPLX && (!has_c || cmax_bits[1] == 0u);
    const double down = ldexp(1.0, -sc);
    for (;;) {
        if (threadIdx.x == 0) chunk_id = atomicAdd(ctrl + 1, 1);
        __syncthreads();
        const int cid = chunk_id;
        if (cid >= total) break;
        const int4 ch = chunks[cid];
        const int b = ch.x;
        const int ox = (b / (nbx * nbx)) * bs, oy = ((b / nbx) % nbx) * bs, oz = (b % nbx) * bs;
        for (int i = threadIdx.x; i < NC * L3; i += kSpreadThreads) acc[i] = (AccT)0;
        if (threadIdx.x < NCL) {
            cls_count[threadIdx.x] = 0;
            cls_fill[threadIdx.x] = 0;
        }
        __syncthreads();
        // deal the chunk's points to lanes by the LDS bank pair of their first accumulator
        // (class = its cell index mod 16; with 32-bit accumulators the bank itself, mod
        // 32): lanes that share a class sit NCL lanes apart, so every atomic of the
        // 343-add sequence, which adds the same offset for all lanes, hits NCL distinct
        // bank pairs (banks) within each group of NCL lanes
        const int npts = kSpreadDeal ? ch.z - ch.y : 0;  // <= kChunk
        for (int i = threadIdx.x; i < npts; i += kSpreadThreads) {
            float t;
            const float4 q = sorted[ch.y + i];
            const int ux = first_node_wrapped(q.x, hi, lo, n, W, &t) - ox;
            const int uy = first_node_wrapped(q.y, hi, lo, n, W, &t) - oy;
            const int uz = first_node_wrapped(q.z, hi, lo, n, W, &t) - oz;
            // the cell index, not the accumulator index: a complex cell's two accumulators
            // would give only even classes and leave half the lanes without points
            // (complex type 1 over 41 M points: 20.1 ms that way, 11.9 ms this way)
            const int cl = (ux * SSP + uy * SP + uz) & (NCL - 1);
            cls_of[i] = (uint8_t)cl;
            atomicAdd(cls_count + cl, 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int s0 = 0, mx = 0;
            for (int q = 0; q < NCL; ++q) {
                cls_start[q] = s0;
                s0 += cls_count[q];
                mx = max(mx, cls_count[q]);
            }
            rounds = (mx + kSpreadThreads / NCL - 1) / (kSpreadThreads / NCL);
        }
        __syncthreads();
        // within a class, arrival rank r goes to position r K mod count (K = 2654435761,
        // a prime above any count, so a bijection): the lanes of one class that add in
        // the same round get points far apart in the chunk's order.  Chunk order follows
        // the input, and a solve's input is orientation-major, so neighbours in it are
        // pixels of one detector patch whose atomics hit the same cells at once (real
        // strengths bound at setpts, 81.92 M image points at M = 81: 6.14 -> 5.67-5.69 ms;
        // the sums are integers, so the result does not change)
        for (int i = threadIdx.x; i < npts; i += kSpreadThreads) {
            const int cl = cls_of[i];
            const int rk = atomicAdd(cls_fill + cl, 1);
            const unsigned cn = (unsigned)cls_count[cl];
            dealt[cls_start[cl] + (int)(((unsigned)rk * (2654435761u % cn)) % cn)] = (uint16_t)i;
        }
        __syncthreads();
        const int my_cls = threadIdx.x & (NCL - 1), my_sub = threadIdx.x / NCL;
        const int nrounds = kSpreadDeal ? rounds : (ch.z - ch.y + kSpreadThreads - 1) / kSpreadThreads;
        // waves in the atomic rounds issue before the other workgroups' dealing and flush
        // (unit strengths at 162 modes over 81.92 M points: 4.62 -> 4.53 ms)
        __builtin_amdgcn_s_setprio(1);
        for (int r = 0; r < nrounds; ++r) {
            int k;
            if (kSpreadDeal) {
                const int e = r * (kSpreadThreads / NCL) + my_sub;
                if (e >= cls_count[my_cls]) continue;
                k = ch.y + dealt[cls_start[my_cls] + e];
            } else {
                k = ch.y + r * kSpreadThreads + threadIdx.x;
                if (k >= ch.z) continue;
            }
            float tx, ty, tz;
            const float4 q = sorted[k];
            const int src = __float_as_int(q.w);
            const int lxw = first_node_wrapped(q.x, hi, lo, n, W, &tx);
            const int lyw = first_node_wrapped(q.y, hi, lo, n, W, &ty);
            const int lzw = first_node_wrapped(q.z, hi, lo, n, W, &tz);
            int lx, ly, lz;
            float wx[W], wy[W], wz[W];
            local_weights<W>(tx, lxw, ox, esk, &lx, wx);
            local_weights<W>(ty, lyw, oy, esk, &ly, wy);
            local_weights<W>(tz, lzw, oz, esk, &lz, wz);
            // lx, ly, lz lie in [0, bs) by construction of the bins; never index outside LDS
            if ((unsigned)lx >= (unsigned)bs || (unsigned)ly >= (unsigned)bs || (unsigned)lz >= (unsigned)bs) continue;
            float vr, vi = 0.f;
            if (CPLX) {
                const float2 cv = reinterpret_cast<const float2*>(c)[src];
                vr = ldexpf(cv.x, sc);
                vi = ldexpf(cv.y, sc);
            } else {
                vr = U32 ? (float)(1 << kOnes32Scale)
                         : ldexpf(bound ? q.w : (c ? reinterpret_cast<const float*>(c)[src] : 1.0f), sc);
            }
            AccT* base = acc + NC * (lx * SSP + ly * SP + lz);
            // rounding offsets of the signed conversion: every tap of the point has the sign
            // of its strength (the weights are >= 0)
            const float hr = copysignf(0.5f, vr), hi2 = copysignf(0.5f, vi);
            if (U32) {
#pragma unroll
                for (int a = 0; a < W; ++a)
#pragma unroll
                    for (int bb = 0; bb < W; ++bb) {
                        AccT* row = base + (a * SSP + bb * SP);
                        const float pr = (wx[a] * wy[bb]) * vr;
#pragma unroll
                        for (int cc = 0; cc < W; ++cc) atomicAdd(row + cc, (unsigned)fmaf(wz[cc], pr, 0.5f));
                    }
            } else if (nonneg) {
                // real non-negative strengths (image counts, unit strengths): every
                // contribution is >= 0, so the cheaper unsigned conversion applies
#pragma unroll
                for (int a = 0; a < W; ++a)
#pragma unroll
                    for (int bb = 0; bb < W; ++bb) {
                        AccT* row = base + (a * SSP + bb * SP);
                        const float pr = (wx[a] * wy[bb]) * vr;
#pragma unroll
                        for (int cc = 0; cc < W; ++cc) atomicAdd(row + cc, (AccT)fixed_addend<true>(wz[cc], pr, 0.5f));
                    }
            } else {
#pragma unroll
                for (int a = 0; a < W; ++a)
#pragma unroll
                    for (int bb = 0; bb < W; ++bb) {
                        const float wab = wx[a] * wy[bb];
                        AccT* row = base + NC * (a * SSP + bb * SP);
                        const float pr = wab * vr;
                        const float pi = wab * vi;
#pragma unroll
                        for (int cc = 0; cc < W; ++cc) {
                            atomicAdd(row + NC * cc, (AccT)fixed_addend<false>(wz[cc], pr, hr));
                            if (CPLX) atomicAdd(row + NC * cc + 1, (AccT)fixed_addend<false>(wz[cc], pi, hi2));
                        }
                    }
            }
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        for (int i = threadIdx.x; i < S3; i += kSpreadThreads) {
            const int ix = i / (S * S), iy = (i / S) % S, iz = i % S;
            const int li = ix * SSP + iy * SP + iz;
            int gx = ox + ix, gy = oy + iy, gz = oz + iz;
            gx = gx >= n ? gx - n : gx;
            gy = gy >= n ? gy - n : gy;
            gz = gz >= n ? gz - n : gz;
            const uint32_t off = (uint32_t)((gx * n + gy) * n + gz);
            if (!CPLX && fixed) {
                // U32: the 2^kOnes32Scale sums moved to the grid's scale 2^sc (sc >= kOnes32Scale)
                const unsigned long long v = U32 ? ((unsigned long long)acc[li] << (sc - kOnes32Scale))
                                                 : (unsigned long long)acc[li];
                if (v != 0ull) atomicAdd(fixed + off, v);
                continue;
            }
            float* dst = reinterpret_cast<float*>(grid) + (size_t)NC * off;
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                const long long v = (long long)acc[NC * li + q];
                if (v != 0) unsafeAtomicAdd(dst + q, (float)((double)v * down));
            }
        }
        __syncthreads();
    }
}

// Fixed-point fine grid of a real spread to float: value = int64 sum x 2^-(sc - shift), the
// scale the spread used (from the same max |c| in device memory).
__global__ void fixed_to_float_kernel(const unsigned long long* __restrict__ fixed, int64_t count,
                                      const unsigned* __restrict__ cmax_bits, int has_c, int shift,
                                      float* __restrict__ out) {
    const int sc = spread_scale_exp(has_c ? __uint_as_float(*cmax_bits) : 1.0f) - shift;
    const double down = ldexp(1.0, -sc);
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += (int64_t)gridDim.x * kBlock)
        out[i] = (float)((double)(long long)fixed[i] * down);
}

// Sorted point k of the interpolation: coordinates and output index.  SortedCopies:
// the sorted coordinate copies of setpts; SortedRotated (below): regenerated from
// perm[k] for the slicing route.
struct SortedCopies {
    const float4* sorted;
    __device__ __forceinline__ int64_t at(int k, float* px, float* py, float* pz) const {
        const float4 q = sorted[k];
        *px = q.x;
        *py = q.y;
        *pz = q.z;
        return __float_as_int(q.w);
    }
};

// Real interpolation: out[j * ostride] = sum of the w^3 real-weighted taps of the
// (real) fine grid.  A complex result is two such passes, over the real and the
// imaginary part of the fine grid, into the interleaved output (ostride 2).
template <int W, class Sorted>
__global__ void __launch_bounds__(kSpreadBlock) interp_binned_kernel(
    const float* __restrict__ grid, int n, const EsKernel esk, float hi, float lo, int bs, int nbx, Sorted pts,
    float* __restrict__ out, int ostride, const int4* __restrict__ chunks, int* __restrict__ ctrl) {
    extern __shared__ float4 sub_raw[];  // 16-byte aligned for the wide row reads
    float* sub = reinterpret_cast<float*>(sub_raw);
    // LDS layout x * SSP + y * SP + z (interp_pitch)
    const int S = bs + W - 1, S3 = S * S * S;
    int SP, SSP;
    interp_pitch(S, W, &SP, &SSP);
    __shared__ int chunk_id;
    const int total = ctrl[0];
    for (;;) {
        if (threadIdx.x == 0) chunk_id = atomicAdd(ctrl + 1, 1);
        __syncthreads();
        const int cid = chunk_id;
        if (cid >= total) break;
        const int4 ch = chunks[cid];
        const int b = ch.x;
        const int ox = (b / (nbx * nbx)) * bs, oy = ((b / nbx) % nbx) * bs, oz = (b % nbx) * bs;
        for (int i = threadIdx.x; i < S3; i += kSpreadBlock) {
            const int ix = i / (S * S), iy = (i / S) % S, iz = i % S;
            int gx = ox + ix, gy = oy + iy, gz = oz + iz;
            gx = gx >= n ? gx - n : gx;
            gy = gy >= n ? gy - n : gy;
            gz = gz >= n ? gz - n : gz;
            sub[ix * SSP + iy * SP + iz] = grid[(uint32_t)((gx * n + gy) * n + gz)];
        }
        __syncthreads();
        for (int k = ch.y + threadIdx.x; k < ch.z; k += kSpreadBlock) {
            float tx, ty, tz, px, py, pz;
            const int64_t j = pts.at(k, &px, &py, &pz);
            const int lxw = first_node_wrapped(px, hi, lo, n, W, &tx);
            const int lyw = first_node_wrapped(py, hi, lo, n, W, &ty);
            const int lzw = first_node_wrapped(pz, hi, lo, n, W, &tz);
            int lx, ly, lz;
            float wx[W], wy[W], wz[W];
            local_weights<W>(tx, lxw, ox, esk, &lx, wx);
            local_weights<W>(ty, lyw, oy, esk, &ly, wy);
            local_weights<W>(tz, lzw, oz, esk, &lz, wz);
            // lx, ly, lz lie in [0, bs) by construction of the bins; never index outside LDS
            if ((unsigned)lx >= (unsigned)bs || (unsigned)ly >= (unsigned)bs || (unsigned)lz >= (unsigned)bs) {
                lx = 0;
                ly = 0;
                lz = 0;
            }
            const int boff = lx * SSP + ly * SP + lz;
            const WideRow<W, CMTIP_INTERP_READS> wr(boff);
            float wl[W];
            weights_at<W>(tz, wr, esk, wl);
            const float* base = sub + boff;
            float sr = 0.f;
#pragma unroll
            for (int a = 0; a < W; ++a) {
                float pr = 0.f;
#pragma unroll
                for (int bb = 0; bb < W; ++bb) pr = fmaf(wy[bb], wr.dot(base + a * SSP + bb * SP, wl), pr);
                sr = fmaf(wx[a], pr, sr);
            }
            out[j * ostride] = sr;
        }
        __syncthreads();
    }
}

// Slice moments in a fixed order: every tile (or row of 16 lanes of a tile, or wave)
// leaves its (sum, sum of squares) in its own slot, stage 1 sums contiguous slot ranges
// per block with a fixed tree, stage 2 adds the block sums in order, so the moments do
// not depend on scheduling.
__global__ void __launch_bounds__(kBlock) moments_stage1_kernel(const double* __restrict__ part, int64_t nslots,
                                                                double* __restrict__ stage) {
    __shared__ double red[2][kBlock];
    const int64_t per = (nslots + kMomentBlocks - 1) / kMomentBlocks;
    const int64_t b0 = (int64_t)blockIdx.x * per, b1 = min(nslots, b0 + per);
    double a = 0.0, q = 0.0;
    for (int64_t i = b0 + threadIdx.x; i < b1; i += kBlock) {
        a += part[2 * i];
        q += part[2 * i + 1];
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = q;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            red[0][threadIdx.x] += red[0][threadIdx.x + h];
            red[1][threadIdx.x] += red[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        stage[2 * blockIdx.x] = red[0][0];
        stage[2 * blockIdx.x + 1] = red[1][0];
    }
}

__global__ void moments_stage2_kernel(const double* __restrict__ stage, double* __restrict__ moments) {
    if (threadIdx.x == 0) {
        double a = 0.0, q = 0.0;
        for (int i = 0; i < kMomentBlocks; ++i) {
            a += stage[2 * i];
            q += stage[2 * i + 1];
        }
        moments[0] += a;
        moments[1] += q;
    }
}

__global__ void fill_f64_kernel(double* __restrict__ p, int64_t count, double v) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < count) p[i] = v;
}

__global__ void fill_i32_kernel(int* __restrict__ p, int64_t count, int v) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t k = i; k < count; k += stride) p[k] = v;
}

__global__ void fill_f32_kernel(float* __restrict__ p, int64_t count, float v) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t k = i; k < count; k += stride) p[k] = v;
}

// ---------------------------------------------------------------- rotations

// Inverse of the rotation matrix of a scalar-first quaternion.  The matrix entries
// follow cmtip/eval/geom.py:25-45 (quaternion products and sums in the quaternion's
// float32 precision, the 1 - 2(.) and 2(.) steps in float64), and the inverse is the
// exact adjugate / determinant in float64, as np.linalg.inv on that matrix.
__device__ void quat_rotinv(const float* q, double* out) {
#pragma clang fp contract(off)
    const float q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const float q01 = q0 * q1, q02 = q0 * q2, q03 = q0 * q3;
    const float q11 = q1 * q1, q12 = q1 * q2, q13 = q1 * q3;
    const float q22 = q2 * q2, q23 = q2 * q3, q33 = q3 * q3;
    double R[9];
    R[0] = 1.0 - 2.0 * (double)(q22 + q33);
    R[1] = 2.0 * (double)(q12 - q03);
    R[2] = 2.0 * (double)(q13 + q02);
    R[3] = 2.0 * (double)(q12 + q03);
    R[4] = 1.0 - 2.0 * (double)(q11 + q33);
    R[5] = 2.0 * (double)(q23 - q01);
    R[6] = 2.0 * (double)(q13 - q02);
    R[7] = 2.0 * (double)(q23 + q01);
    R[8] = 1.0 - 2.0 * (double)(q11 + q22);
    const double c00 = R[4] * R[8] - R[5] * R[7];
    const double c01 = R[5] * R[6] - R[3] * R[8];
    const double c02 = R[3] * R[7] - R[4] * R[6];
    const double det = R[0] * c00 + R[1] * c01 + R[2] * c02;
    const double id = 1.0 / det;
    out[0] = c00 * id;
    out[1] = (R[2] * R[7] - R[1] * R[8]) * id;
    out[2] = (R[1] * R[5] - R[2] * R[4]) * id;
    out[3] = c01 * id;
    out[4] = (R[0] * R[8] - R[2] * R[6]) * id;
    out[5] = (R[2] * R[3] - R[0] * R[5]) * id;
    out[6] = c02 * id;
    out[7] = (R[1] * R[6] - R[0] * R[7]) * id;
    out[8] = (R[0] * R[4] - R[1] * R[3]) * id;
}

__global__ void rotinv_kernel(const float* __restrict__ quats, int64_t n, double* __restrict__ rinv) {
    const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (o >= n) return;
    quat_rotinv(quats + 4 * o, rinv + 9 * o);
}

// x = fl32(fl32(row . s) / ext) * fl32(pi): the reference's einsum in float64, then
// astype(float32), / reciprocal_extent, * np.pi in float32 (align.py:52-57).  iext =
// 1 / ext rounded to double (inv_ext); iext == 0 (ext == 0) returns the unscaled
// fl32(row . s) (autocorrelation.py:27,210).  The float quotient hf / ext is formed as
// fl32(fl64(hf * iext)): that product is within 2^-52 (relative) of the exact quotient,
// and a quotient of two floats is either representable or farther than 2^-49 from
// every float rounding midpoint (hf = H 2^a, ext = E 2^b, a midpoint m = K 2^c with 24-,
// 24- and 25-bit integers H, E and odd K: hf - m ext is a nonzero multiple of
// 2^min(a, b + c)), so rounding the product to float gives the IEEE quotient bit for
// bit, for three instructions instead of a division sequence.
__device__ __forceinline__ float rot_coord(const double* r, double s0, double s1, double s2, double iext) {
#pragma clang fp contract(off)
    const double h = r[0] * s0 + r[1] * s1 + r[2] * s2;
    const float hf = (float)h;
    return iext == 0.0 ? hf : (float)((double)hf * iext) * 3.14159265358979323846f;
}

// 1 / ext in double for rot_coord (0 for ext == 0: no scaling)
inline double inv_ext(float ext) { return ext == 0.0f ? 0.0 : 1.0 / (double)ext; }

struct PtsRotated {
    const double* rinv;  // 9 doubles per orientation
    const double* s;     // 3 x P
    const int* order;    // P pixel indices or NULL
    int64_t P;
    double iext;         // inv_ext(reciprocal extent)
    // coordinates of point j (j < 2^31: the route works in batches); returns its output
    // index o * P + pixel
    __device__ __forceinline__ int64_t at(int64_t j, float* px, float* py, float* pz) const {
        const uint32_t ju = (uint32_t)j, Pu = (uint32_t)P;
        const uint32_t o = ju / Pu, slot = ju - o * Pu;
        const int64_t p = order ? (int64_t)order[slot] : (int64_t)slot;
        const double* r = rinv + 9 * o;
        const double s0 = s[p], s1 = s[P + p], s2 = s[2 * P + p];
        *px = rot_coord(r, s0, s1, s2, iext);
        *py = rot_coord(r + 3, s0, s1, s2, iext);
        *pz = rot_coord(r + 6, s0, s1, s2, iext);
        return (int64_t)o * P + p;
    }
};

struct SortedRotated {
    PtsRotated rot;
    const int* perm;
    __device__ __forceinline__ int64_t at(int k, float* px, float* py, float* pz) const {
        return rot.at(perm[k], px, py, pz);
    }
};

__global__ void __launch_bounds__(kBlock) rotate_coords_kernel(const double* __restrict__ rinv,
                                                               const double* __restrict__ s, int64_t P,
                                                               int64_t total, double iext,
                                                               float* __restrict__ x,
                                                               float* __restrict__ y,
                                                               float* __restrict__ z) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= total) return;
    const int64_t o = j / P, p = j - o * P;
    const double* r = rinv + 9 * o;
    const double s0 = s[p], s1 = s[P + p], s2 = s[2 * P + p];
    x[j] = rot_coord(r, s0, s1, s2, iext);
    y[j] = rot_coord(r + 3, s0, s1, s2, iext);
    z[j] = rot_coord(r + 6, s0, s1, s2, iext);
}

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Moments (striped partials) and the optional bf16 copy of slices written by the
// binned interpolation route.
__global__ void __launch_bounds__(kBlock) slice_finish_kernel(const float* __restrict__ out, int64_t count,
                                                              uint16_t* __restrict__ out_bf16,
                                                              double* __restrict__ part) {
    double s1v = 0.0, s2v = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < count; j += stride) {
        const float v = out[j];
        if (out_bf16) out_bf16[j] = f32_to_bf16(v);
        s1v += (double)v;
        s2v += (double)v * (double)v;
    }
    if (!part) return;
    for (int off = 32; off > 0; off >>= 1) {
        s1v += __shfl_down(s1v, off, 64);
        s2v += __shfl_down(s2v, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        double* slot = part + 2 * ((int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
        slot[0] = s1v;
        slot[1] = s2v;
    }
}

// ---------------------------------------------------------------- fused slicing
//
// Each wave takes one tile of 64 pixels of one orientation.  With the pixels in a
// spatially compact order (8 x 8 detector blocks, see cmtip_slice_real's order
// argument), the tile's rotated points occupy a small box of the fine grid: the wave
// stages that box (its lanes' kernel footprints, at most kBoxFloats floats) in LDS
// with coalesced loads along z, then every lane interpolates from LDS.  Tiles whose
// box does not fit (rare: very tilted, strongly curved or scattered pixel sets) read
// the fine grid directly.  Workgroups are renumbered so that consecutive tiles of an
// orientation run on the same XCD and share its L2.

#ifndef CMTIP_SLICE_WAVES
#define CMTIP_SLICE_WAVES 4
#endif
#ifndef CMTIP_BOX_FLOATS
#define CMTIP_BOX_FLOATS 3328
#endif
constexpr int kSliceBlock = 64 * CMTIP_SLICE_WAVES;  // waves (tiles) per workgroup
constexpr int kBoxFloats = CMTIP_BOX_FLOATS;         // per wave: 52 KB per workgroup, three per CU
constexpr int kBoxGroups = 6;   // (y, z) pairs of an x plane: at most 64 * kBoxGroups
// tile-binned route (see slice_tiles_kernel)
constexpr int kTileS = 24;                    // subgrid edge (cells)
constexpr int kTileE = 19;                    // widest keyed footprint per axis (kernel width and slack included)
constexpr int kTileBs = kTileS - kTileE + 1;  // bin edge (cells)
#ifndef CMTIP_TILE_SP_PAD
#define CMTIP_TILE_SP_PAD 4
#endif
#ifndef CMTIP_TILE_SSP_PAD
#define CMTIP_TILE_SSP_PAD 12
#endif
// Row / plane pitch of the tile route's subgrid: multiples of the wide-row alignment,
// padded so that neighbouring pixels of a tile start their rows in different banks
// (row pitch 28, plane pitch 684 = 12 mod 32; measured at M = 81 over 163.8 M pixels:
// plane pads 4 / 8 / 12 / 16 / 20 give 6.86 / 6.06 / 6.00 / 5.95 / 5.97 ms, row pitch
// 36 gives 7.4 ms).
template <int W>
__host__ __device__ inline void tile_pitch(int* SP, int* SSP) {
    const int A = (W == 7 && CMTIP_INTERP_READS == 4) ? 4 : (CMTIP_INTERP_READS == 1 ? 1 : 2);
    *SP = (kTileS + A - 1) / A * A + CMTIP_TILE_SP_PAD;
    *SSP = kTileS * *SP + CMTIP_TILE_SSP_PAD;
}
#ifndef CMTIP_TILE_CHUNK
#define CMTIP_TILE_CHUNK 128
#endif
constexpr int kTileChunk = CMTIP_TILE_CHUNK;  // tiles per unit of work
constexpr int kTileBlock = 512;               // eight waves: one tile each at a time
constexpr int kStageRows = kTileBlock / kTileS;  // subgrid rows staged per pass
constexpr int kBoxPlanes = 4;   // x planes loaded per step (loads in flight per lane)

__device__ __forceinline__ int wave_min_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Workgroup b runs on XCD b % 8; give each XCD a contiguous run of logical blocks.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t x = b % 8u, q = nb / 8u, r = nb % 8u;
    return x * q + (x < r ? x : r) + b / 8u;
}

template <int W>
__device__ __forceinline__ void row_weights(float t, int l0, const EsKernel& K, float* wt) {
    es_taps<W>(t - (float)l0, K, wt);
}

__device__ __forceinline__ int first_node(float t, int W) { return (int)ceilf(t - 0.5f * (float)W); }

// l mod n in [0, n) for the bounded node indices of the slicing kernels (|l| < 2^22)
__device__ __forceinline__ int wrap_n(int l, int n) { return mod_small(l, n); }

// Interpolation of one point straight from the fine grid (no LDS staging): first
// nodes lx, ly, lz (unwrapped), x / y weights given, z weights from tz.
template <int W>
__device__ __forceinline__ float direct_gather(const float* __restrict__ g, int n, float tz, int lx, int ly, int lz,
                                               const float* wx, const float* wy, const EsKernel& K) {
    float wz[W];
    row_weights<W>(tz, lz, K, wz);
    int ix[W], iy[W], iz[W];
    const int bx = wrap_n(lx, n), by = wrap_n(ly, n), bz = wrap_n(lz, n);
#pragma unroll
    for (int a = 0; a < W; ++a) {
        ix[a] = bx + a >= n ? bx + a - n : bx + a;
        iy[a] = by + a >= n ? by + a - n : by + a;
        iz[a] = bz + a >= n ? bz + a - n : bz + a;
    }
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < W; ++a) {
        float pa = 0.f;
        for (int b = 0; b < W; ++b) {
            const float* row = g + (uint32_t)((ix[a] * n + iy[b]) * n);
            float q = 0.f;
#pragma unroll
            for (int c = 0; c < W; ++c) q = fmaf(wz[c], row[iz[c]], q);
            pa = fmaf(wy[b], q, pa);
        }
        acc = fmaf(wx[a], pa, acc);
    }
    return acc;
}

template <int W>
__global__ void __launch_bounds__(kSliceBlock) slice_real_kernel(const float* __restrict__ g, int n, const EsKernel esk,
                                                            float hi, float lo,
                                                            const double* __restrict__ rinv,
                                                            const double* __restrict__ s,
                                                            const int* __restrict__ order, int64_t P,
                                                            int64_t tiles_per_orient, int64_t total_tiles,
                                                            double iext, float* __restrict__ out,
                                                            uint16_t* __restrict__ out_bf16,
                                                            double* __restrict__ moments,
                                                            int* __restrict__ fit_count) {
    __shared__ float4 box_all[kSliceBlock / 64][kBoxFloats / 4];  // 16-byte aligned rows
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float* box = reinterpret_cast<float*>(box_all[wave]);
    const int64_t tile = (int64_t)xcd_block(blockIdx.x, gridDim.x) * (kSliceBlock / 64) + wave;
    const bool tile_ok = tile < total_tiles;
    const int64_t o = tile_ok ? tile / tiles_per_orient : 0;
    const int64_t slot = (tile_ok ? tile - o * tiles_per_orient : 0) * 64 + lane;
    const bool active = tile_ok && slot < P;
    const int64_t p = order ? (int64_t)order[active ? slot : 0] : (active ? slot : 0);
    const double* r = rinv + 9 * o;
    const double s0 = s[p], s1 = s[P + p], s2 = s[2 * P + p];
    const float tx = fine_coord(rot_coord(r, s0, s1, s2, iext), hi, lo, n);
    const float ty = fine_coord(rot_coord(r + 3, s0, s1, s2, iext), hi, lo, n);
    const float tz = fine_coord(rot_coord(r + 6, s0, s1, s2, iext), hi, lo, n);
    const int lx = first_node(tx, W), ly = first_node(ty, W), lz = first_node(tz, W);
    // box of the active lanes' footprints (wave-uniform values held in scalar registers)
    const int big = 1 << 29;
    const int mnx = __builtin_amdgcn_readfirstlane(wave_min_i(active ? lx : big));
    const int mny = __builtin_amdgcn_readfirstlane(wave_min_i(active ? ly : big));
    const int mnz = __builtin_amdgcn_readfirstlane(wave_min_i(active ? lz : big));
    const int ex = __builtin_amdgcn_readfirstlane(wave_max_i(active ? lx : -big) + W - mnx);
    const int ey = __builtin_amdgcn_readfirstlane(wave_max_i(active ? ly : -big) + W - mny);
    const int ez = __builtin_amdgcn_readfirstlane(wave_max_i(active ? lz : -big) + W - mnz);
    const bool any = mnx < big;
    // box layout x * eyzp + y * ezp + z, the z pitch padded to the wide-row alignment
    constexpr int kA = WideRow<W, CMTIP_SLICE_READS>::kAlign;
    const int ezp = (ez + kA - 1) & ~(kA - 1), eyzp = ey * ezp;
    const bool use_lds = any && ex <= n && ey <= n && ez <= n && ey * ez <= 64 * kBoxGroups &&
                         ex * eyzp <= kBoxFloats;
    if (fit_count && any && lane == 0) {
        // route probe: tiles staged in LDS, tiles with pixels, tiles within the tile
        // route's footprint margin
        atomicAdd(fit_count, use_lds ? 1 : 0);
        atomicAdd(fit_count + 1, 1);
        atomicAdd(fit_count + 2, (ex <= kTileE && ey <= kTileE && ez <= kTileE) ? 1 : 0);
    }
    if (use_lds) {
        // lane q of group gi owns the (y, z) pair q = gi * 64 + lane of the box's x planes:
        // its fine-grid and LDS offsets are fixed, so each x plane costs one scalar offset
        const int bx = wrap_n(mnx, n), by = wrap_n(mny, n), bz = wrap_n(mnz, n);
        const int eyz = ey * ez;
        const float inv_z = 1.0f / (float)ez;
        int goff[kBoxGroups], loff[kBoxGroups];
#pragma unroll
        for (int gi = 0; gi < kBoxGroups; ++gi) {
            const int q = gi * 64 + lane;
            const int j = (int)(((float)q + 0.5f) * inv_z);
            const int k = q - j * ez;
            int gy = by + j, gz = bz + k;
            gy = gy >= n ? gy - n : gy;
            gz = gz >= n ? gz - n : gz;
            goff[gi] = q < eyz ? gy * n + gz : -1;
            loff[gi] = j * ezp + k;
        }
        const int nn = n * n;
        for (int i0 = 0; i0 < ex; i0 += kBoxPlanes) {
            float v[kBoxPlanes][kBoxGroups];
#pragma unroll
            for (int di = 0; di < kBoxPlanes; ++di) {
                int gx = bx + i0 + di;
                gx = gx >= n ? gx - n : gx;
                const float* plane = g + (uint32_t)(gx * nn);
#pragma unroll
                for (int gi = 0; gi < kBoxGroups; ++gi)
                    v[di][gi] = (i0 + di < ex && goff[gi] >= 0) ? plane[goff[gi]] : 0.f;
            }
#pragma unroll
            for (int di = 0; di < kBoxPlanes; ++di)
#pragma unroll
                for (int gi = 0; gi < kBoxGroups; ++gi)
                    if (i0 + di < ex && goff[gi] >= 0) box[(i0 + di) * eyzp + loff[gi]] = v[di][gi];
        }
    }
    __syncthreads();
    float val = 0.f;
    if (active) {
        float wx[W], wy[W];
        row_weights<W>(tx, lx, esk, wx);
        row_weights<W>(ty, ly, esk, wy);
        float acc = 0.f;
        if (use_lds) {
            const int boff = (lx - mnx) * eyzp + (ly - mny) * ezp + (lz - mnz);
            const WideRow<W, CMTIP_SLICE_READS> wr(boff);
            float wl[W];
            weights_at<W>(tz, wr, esk, wl);
            const float* base = box + boff;
#pragma unroll
            for (int a = 0; a < W; ++a) {
                float pa = 0.f;
#pragma unroll
                for (int b = 0; b < W; ++b) pa = fmaf(wy[b], wr.dot(base + a * eyzp + b * ezp, wl), pa);
                acc = fmaf(wx[a], pa, acc);
            }
        } else {
            acc = direct_gather<W>(g, n, tz, lx, ly, lz, wx, wy, esk);
        }
        val = acc;
        const int64_t j = o * P + p;
        out[j] = val;
        if (out_bf16) out_bf16[j] = f32_to_bf16(val);
    }
    if (moments) {
        const double v = active ? (double)val : 0.0;
        double s1v = v, s2v = v * v;
        for (int off = 32; off > 0; off >>= 1) {
            s1v += __shfl_down(s1v, off, 64);
            s2v += __shfl_down(s2v, off, 64);
        }
        if (lane == 0 && tile_ok) {
            moments[2 * tile] = s1v;
            moments[2 * tile + 1] = s2v;
        }
    }
}

// ---------------------------------------------------------------- tile-binned slicing
//
// A tile is 64 consecutive pixels of the pixel order (an 8 x 8 detector block) of one
// orientation.  tile_key_kernel sorts tiles by the bin (kTileBs^3 fine cells) of their
// footprint's low corner, estimated from the block's four corner pixels (slots 0, 21,
// 42, 63 of its Morton order) with one cell of slack on each side: the rotated pixel
// coordinates are affine across a block up to the Ewald curvature, far below a cell.
// A footprint keyed no wider than kTileE cells then lies inside its bin's subgrid of
// kTileS^3 cells.  slice_tiles_kernel stages a bin's subgrid in LDS once per chunk of
// up to kTileChunk tiles and interpolates all of them from it with wide row reads, so
// the staging cost is shared by many tiles of many orientations.  It checks every
// tile's exact footprint; tiles that do not fit (and wider tiles, keyed to an overflow
// bin) read the fine grid directly.

struct TileRot {
    const double* rinv;
    const double* s;
    const int* order;
    int64_t P, tpo;  // pixels, tiles per orientation
    double iext;     // inv_ext(reciprocal extent)
    // pixel of lane `lane` in tile t: returns false past the last pixel
    __device__ __forceinline__ bool pixel(int64_t t, int lane, int64_t* o, int64_t* p) const {
        *o = t / tpo;
        const int64_t slot = (t - *o * tpo) * 64 + lane;
        const bool ok = slot < P;
        *p = order ? (int64_t)order[ok ? slot : 0] : (ok ? slot : 0);
        return ok;
    }
    __device__ __forceinline__ void coords(int64_t o, int64_t p, float hi, float lo, int n, float* tx, float* ty,
                                           float* tz) const {
        const double* r = rinv + 9 * o;
        const double s0 = s[p], s1 = s[P + p], s2 = s[2 * P + p];
        *tx = fine_coord(rot_coord(r, s0, s1, s2, iext), hi, lo, n);
        *ty = fine_coord(rot_coord(r + 3, s0, s1, s2, iext), hi, lo, n);
        *tz = fine_coord(rot_coord(r + 6, s0, s1, s2, iext), hi, lo, n);
    }
};

// Bin key of tile t (see above); nbins = the overflow bin.
template <int W>
__device__ __forceinline__ int tile_key(const TileRot& rot, int64_t t, int n, float hi, float lo, int nbx) {
    const int64_t o = t / rot.tpo, first = (t - o * rot.tpo) * 64;
    const int last = (int)min((int64_t)63, rot.P - 1 - first);
    const int corner[4] = {0, 21, 42, 63};
    int mn[3] = {1 << 29, 1 << 29, 1 << 29}, mx[3] = {-(1 << 29), -(1 << 29), -(1 << 29)};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int64_t slot = first + min(corner[c], last);
        const int64_t p = rot.order ? (int64_t)rot.order[slot] : slot;
        float tc[3];
        rot.coords(o, p, hi, lo, n, tc, tc + 1, tc + 2);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int l = first_node(tc[d], W);
            mn[d] = min(mn[d], l);
            mx[d] = max(mx[d], l);
        }
    }
    bool fit = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        mn[d] -= 1;  // slack
        const int ext = mx[d] + 1 - mn[d] + W;
        fit = fit && ext <= kTileE && ext <= n;
    }
    return fit ? ((wrap_n(mn[0], n) / kTileBs) * nbx + wrap_n(mn[1], n) / kTileBs) * nbx + wrap_n(mn[2], n) / kTileBs
               : nbx * nbx * nbx;
}

template <int W>
__global__ void __launch_bounds__(kBlock) tile_key_kernel(TileRot rot, int64_t ntiles, int n, float hi, float lo,
                                                          int nbx, int* __restrict__ tile_bin,
                                                          int* __restrict__ count) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= ntiles) return;
    const int key = tile_key<W>(rot, t, n, hi, lo, nbx);
    tile_bin[t] = key;
    atomicAdd(count + key, 1);
}

// Workgroup-privatised form (the bins, overflow included, fit an LDS histogram).
template <int W>
__global__ void __launch_bounds__(kBinBlock) tile_key_local_kernel(TileRot rot, int64_t ntiles, int64_t per_block,
                                                                  int n, float hi, float lo, int nbx, int nkeys,
                                                                  int* __restrict__ tile_bin,
                                                                  int* __restrict__ count) {
    extern __shared__ int hist[];
    for (int i = threadIdx.x; i < nkeys; i += kBinBlock) hist[i] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * per_block, t1 = min(ntiles, t0 + per_block);
    for (int64_t t = t0 + threadIdx.x; t < t1; t += kBinBlock) {
        const int key = tile_key<W>(rot, t, n, hi, lo, nbx);
        tile_bin[t] = key;
        atomicAdd(hist + key, 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nkeys; i += kBinBlock)
        if (hist[i]) atomicAdd(count + i, hist[i]);
}

// Neighbouring tiles mostly lie in different bins: plain atomics (the order of tiles
// inside a bin does not change any result).
__global__ void tile_scatter_kernel(int64_t ntiles, const int* __restrict__ tile_bin, const int* __restrict__ offset,
                                    int* __restrict__ cursor, int* __restrict__ tperm) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= ntiles) return;
    const int b = tile_bin[t];
    tperm[offset[b] + atomicAdd(cursor + b, 1)] = (int)t;
}

// Sum of v over each row of 16 lanes, left in the row's last lane (15, 31, 47, 63):
// an inclusive scan by DPP row shifts of 1, 2, 4 and 8 applied to both 32-bit halves
// (lanes shifted in from outside the row read zero).  VALU only, no LDS traffic; the
// whole wave must be active.
template <int SH>
__device__ __forceinline__ double row_shr_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x110 | SH, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x110 | SH, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double row16_sum(double v) {
    v += row_shr_f64<1>(v);
    v += row_shr_f64<2>(v);
    v += row_shr_f64<4>(v);
    v += row_shr_f64<8>(v);
    return v;
}

template <int W>
__global__ void __launch_bounds__(kTileBlock) slice_tiles_kernel(
    const float* __restrict__ g, int n, const EsKernel esk, float hi, float lo, TileRot rot, int nbx,
    const int* __restrict__ tperm, const int4* __restrict__ chunks, int* __restrict__ ctrl, float* __restrict__ out,
    uint16_t* __restrict__ out_bf16, double* __restrict__ part) {
    extern __shared__ float4 sub4[];  // 16-byte aligned rows
    float* sub = reinterpret_cast<float*>(sub4);
    int SP, SSP;
    tile_pitch<W>(&SP, &SSP);
    __shared__ int chunk_id;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int total = ctrl[0];
    const int nbins = nbx * nbx * nbx;
    for (;;) {
        if (threadIdx.x == 0) chunk_id = atomicAdd(ctrl + 1, 1);
        __syncthreads();
        const int cid = chunk_id;
        if (cid >= total) break;
        const int4 ch = chunks[cid];
        const bool staged = ch.x < nbins;
        int ox = 0, oy = 0, oz = 0;
        if (staged) {
            ox = (ch.x / (nbx * nbx)) * kTileBs;
            oy = ((ch.x / nbx) % nbx) * kTileBs;
            oz = (ch.x % nbx) * kTileBs;
            // thread (iz, j) copies z-cell iz of the rows j, j + kStageRows, ... of the
            // kTileS^2 (x, y) rows; indices wrap mod n (one subtraction when n >= kTileS)
            const int iz = threadIdx.x % kTileS, j = threadIdx.x / kTileS;
            if (j < kStageRows) {
                const int gz = (oz + iz) % n;
                if (n >= kTileS) {
#pragma unroll 4
                    for (int r = j; r < kTileS * kTileS; r += kStageRows) {
                        const int ix = r / kTileS, iy = r - ix * kTileS;
                        int gx = ox + ix, gy = oy + iy;
                        gx -= gx >= n ? n : 0;
                        gy -= gy >= n ? n : 0;
                        sub[ix * SSP + iy * SP + iz] = g[(uint32_t)((gx * n + gy) * n + gz)];
                    }
                } else {
                    for (int r = j; r < kTileS * kTileS; r += kStageRows) {
                        const int ix = r / kTileS, iy = r - ix * kTileS;
                        const int gx = (ox + ix) % n, gy = (oy + iy) % n;
                        sub[ix * SSP + iy * SP + iz] = g[(uint32_t)((gx * n + gy) * n + gz)];
                    }
                }
            }
        }
        __syncthreads();
        // interpolating waves issue before the other workgroup's staging loads
        // (10000 orientations at M = 81: 4.73 -> 4.69 ms per slicing call)
        __builtin_amdgcn_s_setprio(1);
        for (int k = ch.y + wave; k < ch.z; k += kTileBlock / 64) {
            int64_t o, p;
            const bool active = rot.pixel(tperm[k], lane, &o, &p);
            float tx, ty, tz;
            rot.coords(o, p, hi, lo, n, &tx, &ty, &tz);
            const int lx = first_node(tx, W), ly = first_node(ty, W), lz = first_node(tz, W);
            // local first nodes in the subgrid, anchored at lane 0 (the tile's first pixel,
            // always present): its offset from the bin origin taken mod n on the scalar
            // unit, the other lanes by their distance to it.  A tile with any footprint
            // outside the subgrid is read from the fine grid instead.
            const int rx = __builtin_amdgcn_readfirstlane(lx), ry = __builtin_amdgcn_readfirstlane(ly);
            const int rz = __builtin_amdgcn_readfirstlane(lz);
            const int d0 = wrap_n(rx - ox, n), d1 = wrap_n(ry - oy, n), d2 = wrap_n(rz - oz, n);
            const int vx = lx - rx + d0, vy = ly - ry + d1, vz = lz - rz + d2;
            const bool lane_fits = vx >= 0 && vy >= 0 && vz >= 0 && vx + W <= kTileS && vy + W <= kTileS &&
                                   vz + W <= kTileS;
            const bool fits = staged && __ballot(active && !lane_fits) == 0;
            const int ux = active ? vx : 0, uy = active ? vy : 0, uz = active ? vz : 0;
            float wx[W], wy[W];
            row_weights<W>(tx, lx, esk, wx);
            row_weights<W>(ty, ly, esk, wy);
            float acc = 0.f;
            if (fits) {
                const int boff = ux * SSP + uy * SP + uz;
                const WideRow<W, CMTIP_INTERP_READS> wr(boff);
                float wl[W];
                weights_at<W>(tz, wr, esk, wl);
                const float* base = sub + boff;
#pragma unroll
                for (int a = 0; a < W; ++a) {
                    float pa = 0.f;
#pragma unroll
                    for (int b = 0; b < W; ++b) pa = fmaf(wy[b], wr.dot(base + a * SSP + b * SP, wl), pa);
                    acc = fmaf(wx[a], pa, acc);
                }
            } else {
                acc = direct_gather<W>(g, n, tz, lx, ly, lz, wx, wy, esk);
            }
            if (active) {
                const int64_t j = o * rot.P + p;
                out[j] = acc;
                if (out_bf16) out_bf16[j] = f32_to_bf16(acc);
            }
            if (part) {
                // this tile's moments, one slot per row of 16 lanes (the tile order inside
                // a bin varies from run to run, the slot of a tile does not); the row sums
                // use DPP lane shifts, not LDS-routed shuffles
                const double v = active ? (double)acc : 0.0;
                const double m1 = row16_sum(v), m2 = row16_sum(v * v);
                if ((lane & 15) == 15) {
                    const int64_t slot = 4 * (int64_t)tperm[k] + (lane >> 4);
                    part[2 * slot] = m1;
                    part[2 * slot + 1] = m2;
                }
            }
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- dispatch helpers

template <template <int> class F, class... Args>
int dispatch_w(int w, Args... args) {
    switch (w) {
#define CMTIP_W(WV) \
    case WV: return F<WV>::run(args...);
        CMTIP_W(2) CMTIP_W(3) CMTIP_W(4) CMTIP_W(5) CMTIP_W(6) CMTIP_W(7) CMTIP_W(8) CMTIP_W(9)
        CMTIP_W(10) CMTIP_W(11) CMTIP_W(12) CMTIP_W(13) CMTIP_W(14) CMTIP_W(15) CMTIP_W(16)
#undef CMTIP_W
        default:
            set_error("unsupported kernel width %d", w);
            return CMTIP_ERR_ARG;
    }
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

int cu_count() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
    }
    return cus;
}

// Raise a kernel's dynamic-LDS limit above the default 64 KB: the most it may ask for
// is the workgroup's LDS less the kernel's own static LDS.
void allow_large_lds(const void* fn) {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, fn) == hipSuccess)
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kMaxLds - a.sharedSizeBytes));
}

int reset_work_counter(cmtip_nufft_plan_s* p, hipStream_t st) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(kBlock), 0, st, p->ctrl.as<int>() + 1, (int64_t)1, 0);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

template <int W>
struct InterpLaunch {
    template <class Sorted>
    static void launch(cmtip_nufft_plan_s* p, Sorted src, float* out, int ostride, hipStream_t st) {
        const size_t lds = subgrid_lds(p->bs, W, false, false);
        // default width: up to 6 workgroups per CU as their subgrids fit (81.92 M image
        // points at M = 81: 2.39-2.61 ms with 4, 2.29-2.36 with 6, same results)
        const int fit = (int)std::max<size_t>(1, std::min<size_t>(6, kMaxLds / (lds + kStaticLds)));
        const unsigned grid = (unsigned)(cu_count() * (p->wpc_req > 0 ? p->wpc_req : fit));
        static bool attr_set = false;
        if (!attr_set && lds > kDefaultLds) {
            allow_large_lds((const void*)interp_binned_kernel<W, Sorted>);
            attr_set = true;
        }
        hipLaunchKernelGGL((interp_binned_kernel<W, Sorted>), dim3(grid), dim3(kSpreadBlock), lds, st,
                           p->fine.as<const float>(), p->n, p->esk, p->tc_hi, p->tc_lo, p->bs, p->nbx, src, out,
                           ostride, p->chunks.as<const int4>(), p->ctrl.as<int>());
    }
    // real interpolation of the fine grid into out[j * ostride]; rot: the slicing
    // route's rotated pixels, else the setpts points
    static int run(cmtip_nufft_plan_s* p, float* out, int ostride, const PtsRotated* rot, hipStream_t st) {
        if (p->npts == 0) return CMTIP_OK;
        CMTIP_TRY(reset_work_counter(p, st));
        if (rot)
            launch(p, SortedRotated{*rot, p->perm.as<const int>()}, out, ostride, st);
        else
            launch(p, SortedCopies{p->sorted.as<const float4>()}, out, ostride, st);
        CMTIP_LAUNCH_CHECK();
        return CMTIP_OK;
    }
};

template <int W>
struct SpreadLaunch {
    // bound: real strengths carried in the point records, max |c| recorded at setpts
    static int run(cmtip_nufft_plan_s* p, bool cplx, const void* c, bool bound, hipStream_t st) {
        if (p->npts == 0) return CMTIP_OK;
        CMTIP_TRY(reset_work_counter(p, st));
        unsigned* cmax = reinterpret_cast<unsigned*>(p->ctrl.as<int>() + 2);
        if (c && !bound) {
            hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(kBlock), 0, st, reinterpret_cast<int*>(cmax),
                               (int64_t)2, 0);
            const int64_t cnt = (cplx ? 2 : 1) * p->npts;
            hipLaunchKernelGGL(absmax_kernel, dim3((unsigned)std::min<int64_t>(4096, (cnt + kBlock - 1) / kBlock)),
                               dim3(kBlock), 0, st, reinterpret_cast<const float*>(c), cnt, cmax);
        }
        // unit strengths: 32-bit LDS accumulators (half the subgrid LDS); their flush shifts
        // left by the grid's scale minus kOnes32Scale, which is >= 0 while npts < 2^43
        const bool ones32 = !cplx && !c && !bound && p->npts < (int64_t(1) << 43);
        const size_t lds = subgrid_lds(p->bs, W, true, cplx) / (ones32 ? 2 : 1);
        const int fit = (int)std::max<size_t>(1, std::min<size_t>(4, kMaxLds / (lds + kSpreadStaticLds)));
        const unsigned grid = (unsigned)(cu_count() * (p->wpc_req > 0 ? p->wpc_req : fit));
        const int4* ch = p->chunks.as<const int4>();
        static bool attr_set[3] = {false, false, false};
        const int form = cplx ? 1 : (ones32 ? 2 : 0);
        if (!attr_set[form] && lds > kDefaultLds) {
            allow_large_lds(form == 1 ? (const void*)spread_binned_kernel<W, true>
                                      : (form == 2 ? (const void*)spread_binned_kernel<W, false, true>
                                                   : (const void*)spread_binned_kernel<W, false>));
            attr_set[form] = true;
        }
        if (cplx) {
            hipLaunchKernelGGL((spread_binned_kernel<W, true>), dim3(grid), dim3(kSpreadThreads), lds, st,
                               p->fine.as<float2>(), p->n, p->esk, p->tc_hi, p->tc_lo, p->bs, p->nbx,
                               p->sorted.as<const float4>(), reinterpret_cast<const float2*>(c), cmax, ch,
                               p->ctrl.as<int>(), 0, (unsigned long long*)nullptr, 0);
            CMTIP_LAUNCH_CHECK();
            return CMTIP_OK;
        }
        // real: the chunks' integer sums go into the zeroed 64-bit grid (exact, so the result
        // does not depend on the order of the adds), scaled so that npts full-weight
        // contributions (each below 2^(30 - shift)) stay below 2^63, then one pass converts
        // it into the float fine grid
        int bits = 0;
        while (bits < 62 && (int64_t(1) << bits) <= p->npts) ++bits;  // npts < 2^bits
        const int shift = std::max(0, bits - 33);
        const int64_t n3 = (int64_t)p->n * p->n * p->n;
        if (ones32)
            hipLaunchKernelGGL((spread_binned_kernel<W, false, true>), dim3(grid), dim3(kSpreadThreads), lds, st,
                               p->fine.as<float>(), p->n, p->esk, p->tc_hi, p->tc_lo, p->bs, p->nbx,
                               p->sorted.as<const float4>(), (const float*)nullptr, cmax, ch, p->ctrl.as<int>(), 0,
                               p->fine64.as<unsigned long long>(), shift);
        else
            hipLaunchKernelGGL((spread_binned_kernel<W, false>), dim3(grid), dim3(kSpreadThreads), lds, st,
                               p->fine.as<float>(), p->n, p->esk, p->tc_hi, p->tc_lo, p->bs, p->nbx,
                               p->sorted.as<const float4>(), reinterpret_cast<const float*>(c), cmax, ch,
                               p->ctrl.as<int>(), bound ? 1 : 0, p->fine64.as<unsigned long long>(), shift);
        hipLaunchKernelGGL(fixed_to_float_kernel, dim3((unsigned)std::min<int64_t>(65536, (n3 + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, st, p->fine64.as<const unsigned long long>(), n3,
                           (const unsigned*)cmax, (c || bound) ? 1 : 0, shift, p->fine.as<float>());
        CMTIP_LAUNCH_CHECK();
        return CMTIP_OK;
    }
};

// Sort the registered points (or, with rot, rotated pixels) into spatial bins (counting
// sort) and build the chunk list.  Registered points and rotated pixels with
// rot_records get 16-byte records (x, y, z, original index or strength); the slicing
// route's rotated pixels (rot without rot_records) get only the permutation.
// cstr (optional): real strengths stored in the records in place of the original
// indices, with their max |c| and sign recorded in ctrl[2], ctrl[3]
int bin_points(cmtip_nufft_plan_s* p, hipStream_t st, const PtsRotated* rot = nullptr,
               const float* cstr = nullptr, bool rot_records = false) {
    const int64_t N = p->npts;
    const bool records = !rot || rot_records;
    p->bs = p->bs_req > 0 ? p->bs_req : bin_size_for(p->w);
    p->chunk = p->chunk_req > 0 ? p->chunk_req : kChunk;
    p->nbx = (p->n + p->bs - 1) / p->bs;
    p->nbins = (int64_t)p->nbx * p->nbx * p->nbx;
    p->max_chunks = p->nbins + N / p->chunk + 1;
    CMTIP_TRY(p->bin_of.ensure(sizeof(int) * (size_t)N));
    CMTIP_TRY(p->count.ensure(sizeof(int) * (size_t)p->nbins));
    CMTIP_TRY(p->offset.ensure(sizeof(int) * (size_t)(p->nbins + 1)));
    CMTIP_TRY(p->nchunk.ensure(sizeof(int) * (size_t)(p->nbins + 1)));
    if (records)
        CMTIP_TRY(p->sorted.ensure(sizeof(float4) * (size_t)N));
    else
        CMTIP_TRY(p->perm.ensure(sizeof(int) * (size_t)N));
    CMTIP_TRY(p->chunks.ensure(sizeof(int4) * (size_t)p->max_chunks));
    CMTIP_TRY(p->ctrl.ensure(sizeof(int) * 4));
    const unsigned gb = (unsigned)((p->nbins + kBlock - 1) / kBlock);
    const unsigned gp = blocks_for(N);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(gb), dim3(kBlock), 0, st, p->count.as<int>(), p->nbins, 0);
    unsigned* cmax = reinterpret_cast<unsigned*>(p->ctrl.as<int>() + 2);
    if (cstr) hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(kBlock), 0, st, p->ctrl.as<int>() + 2, (int64_t)2, 0);
    p->bound = cstr ? 1 : 0;
    // points with records: their max |coordinate| for the caller's range check
    unsigned* amax = nullptr;
    p->amax_valid = 0;
    if (records) {
        CMTIP_TRY(p->amax.ensure(sizeof(unsigned)));
        amax = p->amax.as<unsigned>();
        hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(kBlock), 0, st, reinterpret_cast<int*>(amax), (int64_t)1, 0);
        p->amax_valid = 1;
    }
    const PtsArrays arr{p->x, p->y, p->z};
    if (p->nbins <= kLocalBinsMax) {
        const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((N + 8191) / 8192, 4 * cu_count()));
        const int64_t per_block = (N + blocks - 1) / blocks;
        const size_t lds = sizeof(int) * (size_t)p->nbins;
        const int nb = (int)p->nbins;
        static bool attr_set = false;
        if (!attr_set) {
            allow_large_lds((const void*)bin_count_local_kernel<PtsRotated>);
            allow_large_lds((const void*)bin_count_local_kernel<PtsArrays>);
            allow_large_lds((const void*)bin_scatter_local_kernel<false, PtsArrays>);
            allow_large_lds((const void*)bin_scatter_local_kernel<true, PtsArrays>);
            allow_large_lds((const void*)bin_scatter_local_kernel<true, PtsRotated>);
            attr_set = true;
        }
        if (rot)
            hipLaunchKernelGGL(bin_count_local_kernel<PtsRotated>, dim3((unsigned)blocks), dim3(kBinBlock), lds, st,
                               *rot, N, per_block, p->n, p->w, p->bs, p->nbx, nb, p->tc_hi, p->tc_lo,
                               p->bin_of.as<int>(), p->count.as<int>(), amax, cstr, cmax);
        else
            hipLaunchKernelGGL(bin_count_local_kernel<PtsArrays>, dim3((unsigned)blocks), dim3(kBinBlock), lds, st,
                               arr, N, per_block, p->n, p->w, p->bs, p->nbx, nb, p->tc_hi, p->tc_lo,
                               p->bin_of.as<int>(), p->count.as<int>(), amax, cstr, cmax);
        hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(1024), 0, st, p->count.as<int>(), p->nbins, p->chunk,
                           p->offset.as<int>(), p->nchunk.as<int>(), p->ctrl.as<int>());
        if (!records)
            hipLaunchKernelGGL((bin_scatter_local_kernel<false, PtsArrays>), dim3((unsigned)blocks), dim3(kBinBlock),
                               lds, st, arr, N, per_block, nb, p->bin_of.as<const int>(), p->offset.as<const int>(),
                               p->count.as<int>(), p->perm.as<int>(), (float4*)nullptr, (const float*)nullptr);
        else if (rot)
            hipLaunchKernelGGL((bin_scatter_local_kernel<true, PtsRotated>), dim3((unsigned)blocks), dim3(kBinBlock),
                               lds, st, *rot, N, per_block, nb, p->bin_of.as<const int>(), p->offset.as<const int>(),
                               p->count.as<int>(), (int*)nullptr, p->sorted.as<float4>(), cstr);
        else
            hipLaunchKernelGGL((bin_scatter_local_kernel<true, PtsArrays>), dim3((unsigned)blocks), dim3(kBinBlock),
                               lds, st, arr, N, per_block, nb, p->bin_of.as<const int>(), p->offset.as<const int>(),
                               p->count.as<int>(), (int*)nullptr, p->sorted.as<float4>(), cstr);
        hipLaunchKernelGGL(chunk_fill_kernel, dim3(gb), dim3(kBlock), 0, st, p->offset.as<const int>(),
                           p->nchunk.as<const int>(), p->nbins, p->chunk, p->chunks.as<int4>());
        CMTIP_LAUNCH_CHECK();
        return CMTIP_OK;
    }
    if (rot)
        hipLaunchKernelGGL(bin_count_kernel<PtsRotated>, dim3(gp), dim3(kBlock), 0, st, *rot, N, p->n, p->w, p->bs,
                           p->nbx, p->tc_hi, p->tc_lo, p->bin_of.as<int>(), p->count.as<int>(), amax, cstr, cmax);
    else
        hipLaunchKernelGGL(bin_count_kernel<PtsArrays>, dim3(gp), dim3(kBlock), 0, st, arr, N, p->n, p->w, p->bs,
                           p->nbx, p->tc_hi, p->tc_lo, p->bin_of.as<int>(), p->count.as<int>(), amax, cstr, cmax);
    hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(1024), 0, st, p->count.as<int>(), p->nbins, p->chunk,
                       p->offset.as<int>(), p->nchunk.as<int>(), p->ctrl.as<int>());
    if (!records)
        hipLaunchKernelGGL((bin_scatter_kernel<false, PtsArrays>), dim3(gp), dim3(kBlock), 0, st, arr, N,
                           p->bin_of.as<const int>(), p->offset.as<const int>(), p->count.as<int>(), p->perm.as<int>(),
                           (float4*)nullptr, (const float*)nullptr);
    else if (rot)
        hipLaunchKernelGGL((bin_scatter_kernel<true, PtsRotated>), dim3(gp), dim3(kBlock), 0, st, *rot, N,
                           p->bin_of.as<const int>(), p->offset.as<const int>(), p->count.as<int>(), (int*)nullptr,
                           p->sorted.as<float4>(), cstr);
    else
        hipLaunchKernelGGL((bin_scatter_kernel<true, PtsArrays>), dim3(gp), dim3(kBlock), 0, st, arr, N,
                           p->bin_of.as<const int>(), p->offset.as<const int>(), p->count.as<int>(), (int*)nullptr,
                           p->sorted.as<float4>(), cstr);
    hipLaunchKernelGGL(chunk_fill_kernel, dim3(gb), dim3(kBlock), 0, st, p->offset.as<const int>(),
                       p->nchunk.as<const int>(), p->nbins, p->chunk, p->chunks.as<int4>());
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

int timer_mark(cmtip_nufft_plan_s* p, int which, hipStream_t st);

template <int W>
struct TileLaunch {
    static int run(cmtip_nufft_plan_s* p, const TileRot& rot, int64_t ntiles, float* out, uint16_t* out_bf16,
                   double* part, bool timed, hipStream_t st) {
        const int nbx = (p->n + kTileBs - 1) / kTileBs;
        const int64_t nbins = (int64_t)nbx * nbx * nbx;  // + 1 overflow bin
        CMTIP_TRY(p->bin_of.ensure(sizeof(int) * (size_t)ntiles));
        CMTIP_TRY(p->count.ensure(sizeof(int) * (size_t)(nbins + 1)));
        CMTIP_TRY(p->offset.ensure(sizeof(int) * (size_t)(nbins + 2)));
        CMTIP_TRY(p->nchunk.ensure(sizeof(int) * (size_t)(nbins + 2)));
        CMTIP_TRY(p->perm.ensure(sizeof(int) * (size_t)ntiles));
        CMTIP_TRY(p->chunks.ensure(sizeof(int4) * (size_t)(nbins + 2 + ntiles / kTileChunk)));
        CMTIP_TRY(p->ctrl.ensure(sizeof(int) * 4));
        const unsigned gb = (unsigned)((nbins + 1 + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(fill_i32_kernel, dim3(gb), dim3(kBlock), 0, st, p->count.as<int>(), nbins + 1, 0);
        const bool local = nbins + 1 <= kLocalBinsMax;
        // about two tiles per thread: each tile's corner loads form a dependent chain
        const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((ntiles + 2047) / 2048, 4 * cu_count()));
        const int64_t per_block = (ntiles + blocks - 1) / blocks;
        const size_t hlds = sizeof(int) * (size_t)(nbins + 1);
        if (local) {
            static bool attr_set = false;
            if (!attr_set) {
                allow_large_lds((const void*)tile_key_local_kernel<W>);
                allow_large_lds((const void*)bin_scatter_local_kernel<false, PtsArrays>);
                attr_set = true;
            }
            hipLaunchKernelGGL(tile_key_local_kernel<W>, dim3((unsigned)blocks), dim3(kBinBlock), hlds, st, rot,
                               ntiles, per_block, p->n, p->tc_hi, p->tc_lo, nbx, (int)(nbins + 1),
                               p->bin_of.as<int>(), p->count.as<int>());
        } else {
            hipLaunchKernelGGL(tile_key_kernel<W>, dim3(blocks_for(ntiles)), dim3(kBlock), 0, st, rot, ntiles, p->n,
                               p->tc_hi, p->tc_lo, nbx, p->bin_of.as<int>(), p->count.as<int>());
        }
        hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(1024), 0, st, p->count.as<int>(), nbins + 1, kTileChunk,
                           p->offset.as<int>(), p->nchunk.as<int>(), p->ctrl.as<int>());
        if (local)
            hipLaunchKernelGGL((bin_scatter_local_kernel<false, PtsArrays>), dim3((unsigned)blocks), dim3(kBinBlock),
                               hlds, st, PtsArrays{nullptr, nullptr, nullptr}, ntiles, per_block, (int)(nbins + 1),
                               p->bin_of.as<const int>(), p->offset.as<const int>(), p->count.as<int>(),
                               p->perm.as<int>(), (float4*)nullptr, (const float*)nullptr);
        else
            hipLaunchKernelGGL(tile_scatter_kernel, dim3(blocks_for(ntiles)), dim3(kBlock), 0, st, ntiles,
                               p->bin_of.as<const int>(), p->offset.as<const int>(), p->count.as<int>(),
                               p->perm.as<int>());
        hipLaunchKernelGGL(chunk_fill_kernel, dim3(gb), dim3(kBlock), 0, st, p->offset.as<const int>(),
                           p->nchunk.as<const int>(), nbins + 1, kTileChunk, p->chunks.as<int4>());
        CMTIP_LAUNCH_CHECK();
        CMTIP_TRY(reset_work_counter(p, st));
        int SP, SSP;
        tile_pitch<W>(&SP, &SSP);
        const size_t lds = (size_t)kTileS * SSP * sizeof(float);
        static bool attr_set = false;
        if (!attr_set && lds > kDefaultLds) {
            allow_large_lds((const void*)slice_tiles_kernel<W>);
            attr_set = true;
        }
        const unsigned grid = (unsigned)(cu_count() * std::max<size_t>(1, kMaxLds / (lds + kStaticLds)));
        if (timed) CMTIP_TRY(timer_mark(p, 0, st));
        hipLaunchKernelGGL(slice_tiles_kernel<W>, dim3(grid), dim3(kTileBlock), lds, st, p->fine.as<const float>(),
                           p->n, p->esk, p->tc_hi, p->tc_lo, rot, nbx, p->perm.as<const int>(),
                           p->chunks.as<const int4>(), p->ctrl.as<int>(), out, out_bf16, part);
        CMTIP_LAUNCH_CHECK();
        if (timed) CMTIP_TRY(timer_mark(p, 1, st));
        return CMTIP_OK;
    }
};

template <int W>
struct SliceLaunch {
    static int run(const cmtip_nufft_plan_s* p, const double* s, const int* order, int64_t P, int64_t n_orient,
                   float ext, float* out, uint16_t* out_bf16, double* moments, int* fit_count, hipStream_t st) {
        const int64_t tpo = (P + 63) / 64;
        const int64_t tiles = tpo * n_orient;
        const int64_t blocks = (tiles + kSliceBlock / 64 - 1) / (kSliceBlock / 64);
        if (blocks > 0x7fffffffLL) {
            set_error("slice_real: too many orientations for one launch");
            return CMTIP_ERR_ARG;
        }
        hipLaunchKernelGGL((slice_real_kernel<W>), dim3((unsigned)blocks), dim3(kSliceBlock), 0, st,
                           p->fine.as<const float>(), p->n, p->esk, p->tc_hi, p->tc_lo,
                           p->rinv.as<const double>(), s, order, P, tpo, tiles, inv_ext(ext), out, out_bf16, moments,
                           fit_count);
        CMTIP_LAUNCH_CHECK();
        return CMTIP_OK;
    }
};

int fill_zero(void* p, int64_t nfloats, hipStream_t st) {
    if (nfloats <= 0) return CMTIP_OK;
    int64_t nb = (nfloats + kBlock - 1) / kBlock;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)nb), dim3(kBlock), 0, st, (float*)p, nfloats, 0.0f);
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

// Segments of an n-point line holding M centred modes: LDS positions [0, ceil(M/2))
// hold k >= 0 (array index k + floor(M/2)); [n - floor(M/2), n) hold k < 0
// (array index 0 ...).
void mode_segments_in(AxisDesc& d, int M) {
    d.ia = M - M / 2;
    d.ig1 = M / 2;
    d.ib = M / 2;
    d.ig2 = 0;
}
void mode_segments_out(AxisDesc& d, int M) {
    d.oa = M - M / 2;
    d.og1 = M / 2;
    d.ob = M / 2;
    d.og2 = 0;
}

// Uniform grid f (M^3, float or float2) -> fine grid (n^3), real part only when
// real_out.  Deconvolution weights on every axis; support multiplies f on load.
int grid_to_fine(cmtip_nufft_plan_s* p, const void* f, bool f_real, const float* support, bool real_out,
                 hipStream_t st, bool imag_out = false) {
    const int M = p->M, n = p->n;
    const float* dw = p->deconv.as<const float>();
    // pass 1: axis 2 (contiguous), lines (i0, i1)
    AxisDesc d1;
    d1.in = f;
    d1.in_real = f_real ? 1 : 0;
    d1.in_su = (int64_t)M * M;
    d1.in_sv = M;
    d1.in_sa = 1;
    mode_segments_in(d1, M);
    d1.in_w = dw;
    d1.in_mask = support;
    d1.out = p->tmpB.p;
    d1.out_su = (int64_t)M * n;
    d1.out_sv = n;
    d1.out_sa = 1;
    d1.oa = n;
    d1.og1 = 0;
    d1.nu = M;
    d1.nv = M;
    d1.sign = p->iflag;
    CMTIP_TRY(fft_axis_pass(p->fft, d1, st));
    // pass 2: axis 1, lines (i0, l2)
    AxisDesc d2;
    d2.in = p->tmpB.p;
    d2.in_su = (int64_t)M * n;
    d2.in_sv = 1;
    d2.in_sa = n;
    mode_segments_in(d2, M);
    d2.in_w = dw;
    d2.out = p->tmpA.p;
    d2.out_su = (int64_t)n * n;
    d2.out_sv = 1;
    d2.out_sa = n;
    d2.oa = n;
    d2.og1 = 0;
    d2.nu = M;
    d2.nv = n;
    d2.sign = p->iflag;
    CMTIP_TRY(fft_axis_pass(p->fft, d2, st));
    // pass 3: axis 0, lines (l1, l2)
    AxisDesc d3;
    d3.in = p->tmpA.p;
    d3.in_su = n;
    d3.in_sv = 1;
    d3.in_sa = (int64_t)n * n;
    mode_segments_in(d3, M);
    d3.in_w = dw;
    d3.out = p->fine.p;
    d3.out_mode = imag_out ? 2 : (real_out ? 1 : 0);
    d3.out_su = n;
    d3.out_sv = 1;
    d3.out_sa = (int64_t)n * n;
    d3.oa = n;
    d3.og1 = 0;
    d3.nu = n;
    d3.nv = n;
    d3.sign = p->iflag;
    CMTIP_TRY(fft_axis_pass(p->fft, d3, st));
    return CMTIP_OK;
}

// Fine grid (n^3, float or float2) -> uniform grid f (M^3): crop, deconvolve,
// support, scale, real part when real_out.
int fine_to_grid(cmtip_nufft_plan_s* p, bool fine_real, void* f, bool real_out, const float* support,
                 float scale, hipStream_t st) {
    const int M = p->M, n = p->n;
    const float* dw = p->deconv.as<const float>();
    // pass A: axis 0, lines (l1, l2)
    AxisDesc a;
    a.in = p->fine.p;
    a.in_real = fine_real ? 1 : 0;
    a.in_su = n;
    a.in_sv = 1;
    a.in_sa = (int64_t)n * n;
    a.ia = n;
    a.ig1 = 0;
    a.out = p->tmpA.p;
    a.out_su = n;
    a.out_sv = 1;
    a.out_sa = (int64_t)n * n;
    mode_segments_out(a, M);
    a.out_w = dw;
    a.nu = n;
    a.nv = n;
    a.sign = p->iflag;
    CMTIP_TRY(fft_axis_pass(p->fft, a, st));
    // pass B: axis 1, lines (i0, l2)
    AxisDesc b;
    b.in = p->tmpA.p;
    b.in_su = (int64_t)n * n;
    b.in_sv = 1;
    b.in_sa = n;
    b.ia = n;
    b.ig1 = 0;
    b.out = p->tmpB.p;
    b.out_su = (int64_t)M * n;
    b.out_sv = 1;
    b.out_sa = n;
    mode_segments_out(b, M);
    b.out_w = dw;
    b.nu = M;
    b.nv = n;
    b.sign = p->iflag;
    CMTIP_TRY(fft_axis_pass(p->fft, b, st));
    // pass C: axis 2 (contiguous), lines (i0, i1)
    AxisDesc c;
    c.in = p->tmpB.p;
    c.in_su = (int64_t)M * n;
    c.in_sv = n;
    c.in_sa = 1;
    c.ia = n;
    c.ig1 = 0;
    c.out = f;
    c.out_mode = real_out ? 1 : 0;
    c.out_su = (int64_t)M * M;
    c.out_sv = M;
    c.out_sa = 1;
    mode_segments_out(c, M);
    c.out_w = dw;
    c.out_scale = scale;
    c.out_mask = support;
    c.nu = M;
    c.nv = M;
    c.sign = p->iflag;
    CMTIP_TRY(fft_axis_pass(p->fft, c, st));
    return CMTIP_OK;
}

bool valid_plan(cmtip_nufft_plan p) {
    if (!p) {
        set_error("null plan");
        return false;
    }
    return true;
}

int ensure_rinv(cmtip_nufft_plan_s* p, const float* quats, int64_t n_orient, hipStream_t st) {
    CMTIP_TRY(p->rinv.ensure((size_t)(n_orient > 0 ? n_orient : 1) * 9 * sizeof(double)));
    if (n_orient > 0) {
        hipLaunchKernelGGL(rotinv_kernel, dim3(blocks_for(n_orient)), dim3(kBlock), 0, st, quats, n_orient,
                           p->rinv.as<double>());
        CMTIP_LAUNCH_CHECK();
    }
    return CMTIP_OK;
}

}  // namespace

// ==================================================================== C ABI

extern "C" int cmtip_nufft_makeplan(int type, int n_modes, int iflag, double eps, cmtip_nufft_plan* out) {
    if (!out) {
        set_error("makeplan: null output");
        return CMTIP_ERR_ARG;
    }
    *out = nullptr;
    if ((type != 1 && type != 2) || n_modes < 1 || n_modes > 1024 || (iflag != 1 && iflag != -1)) {
        set_error("makeplan: bad arguments (type=%d, n_modes=%d, iflag=%d)", type, n_modes, iflag);
        return CMTIP_ERR_ARG;
    }
    auto* p = new cmtip_nufft_plan_s();
    p->type = type;
    p->M = n_modes;
    p->iflag = iflag;
    p->eps = eps;
    EsParams es = es_params(eps);
    p->w = es.w;
    p->beta = es.beta;
    p->beta_l2e = (float)(es.beta / M_LN2);
    p->esk.beta_l2e = p->beta_l2e;
    es_poly(p->w, p->beta, p->esk.c);
    p->n = fine_size(n_modes, es.w);
    if ((int64_t)p->n * p->n * p->n >= ((int64_t)1 << 31)) {
        // kernels index the fine grid with 32-bit offsets
        set_error("makeplan: fine grid %d^3 too large (n_modes=%d)", p->n, n_modes);
        delete p;
        return CMTIP_ERR_ARG;
    }
    const double tc = (double)p->n / (2.0 * M_PI);
    p->tc_hi = (float)tc;
    p->tc_lo = (float)(tc - (double)p->tc_hi);
    const int n = p->n, M = n_modes;
    int st = fft1d_init(p->fft, n);
    if (st == CMTIP_OK) st = p->deconv.ensure(sizeof(float) * n);
    if (st == CMTIP_OK) {
        std::vector<float> dw = deconv_weights(M, n, p->w, p->beta);
        if (hipMemcpy(p->deconv.p, dw.data(), sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("makeplan: hipMemcpy failed");
            st = CMTIP_ERR_HIP;
        }
    }
    if (st == CMTIP_OK) st = p->fine.ensure(sizeof(float2) * (size_t)n * n * n);
    if (st == CMTIP_OK) st = p->tmpA.ensure(sizeof(float2) * (size_t)M * n * n);
    if (st == CMTIP_OK) st = p->tmpB.ensure(sizeof(float2) * (size_t)M * M * n);
    if (st != CMTIP_OK) {
        cmtip_nufft_destroy(p);
        return st;
    }
    *out = p;
    return CMTIP_OK;
}

extern "C" int cmtip_nufft_info(cmtip_nufft_plan p, int* w, int* n_fine, double* beta) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (w) *w = p->w;
    if (n_fine) *n_fine = p->n;
    if (beta) *beta = p->beta;
    return CMTIP_OK;
}

extern "C" int cmtip_nufft_set_bins(cmtip_nufft_plan p, int bin_size, int chunk, int wg_per_cu) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    const int bs = bin_size > 0 ? bin_size : bin_size_for(p->w);
    // the subgrid must fit LDS for every execute form of this plan type (complex worst case)
    const size_t static_lds = p->type == 1 ? kSpreadStaticLds : kStaticLds;
    if (bin_size < 0 || bs > p->n || subgrid_lds(bs, p->w, p->type == 1, true) + static_lds > kMaxLds) {
        set_error("set_bins: bin size %d does not fit (w=%d, n=%d)", bin_size, p->w, p->n);
        return CMTIP_ERR_ARG;
    }
    if (chunk < 0 || (chunk > 0 && (chunk < 64 || chunk > kChunk))) {
        set_error("set_bins: chunk %d outside [64, %d]", chunk, kChunk);
        return CMTIP_ERR_ARG;
    }
    if (wg_per_cu < 0 || wg_per_cu > 16) {
        set_error("set_bins: wg_per_cu %d outside [0, 16]", wg_per_cu);
        return CMTIP_ERR_ARG;
    }
    p->bs_req = bin_size;
    p->chunk_req = chunk;
    p->wpc_req = wg_per_cu;
    return CMTIP_OK;
}

extern "C" int cmtip_nufft_get_bins(cmtip_nufft_plan p, int* bin_size, int* chunk, int* wg_per_cu) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (bin_size) *bin_size = p->bs_req > 0 ? p->bs_req : bin_size_for(p->w);
    if (chunk) *chunk = p->chunk_req > 0 ? p->chunk_req : kChunk;
    if (wg_per_cu) *wg_per_cu = p->wpc_req;
    return CMTIP_OK;
}

namespace {
int setpts_impl(cmtip_nufft_plan p, int64_t npts, const float* x, const float* y, const float* z, const float* c,
                bool bind, void* stream) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (npts < 0 || (npts > 0 && (!x || !y || !z))) {
        set_error("setpts: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (npts > 0x7fffffffLL) {
        set_error("setpts: at most 2^31 - 1 points per plan");
        return CMTIP_ERR_ARG;
    }
    p->npts = npts;
    p->x = x;
    p->y = y;
    p->z = z;
    p->bound = 0;
    p->rot_records = 0;
    if (npts == 0) {
        p->amax_valid = 2;  // no points: the maximum is 0
        p->bound = bind ? 1 : 0;
        return CMTIP_OK;
    }
    const int rc = bin_points(p, as_stream(stream), nullptr, c);
    if (rc != CMTIP_OK) p->x = p->y = p->z = nullptr;  // no usable bins: execute refuses
    return rc;
}
}  // namespace

extern "C" int cmtip_nufft_setpts(cmtip_nufft_plan p, int64_t npts, const float* x, const float* y,
                                  const float* z, void* stream) {
    return setpts_impl(p, npts, x, y, z, nullptr, false, stream);
}

extern "C" int cmtip_nufft_setpts_strengths(cmtip_nufft_plan p, int64_t npts, const float* x, const float* y,
                                            const float* z, const float* c, void* stream) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (p->type != 1) {
        set_error("setpts_strengths: type-1 plans only");
        return CMTIP_ERR_ARG;
    }
    if (npts > 0 && !c) {
        set_error("setpts_strengths: null strengths");
        return CMTIP_ERR_ARG;
    }
    return setpts_impl(p, npts, x, y, z, c, true, stream);
}

extern "C" int cmtip_nufft_setpts_rotated(cmtip_nufft_plan p, const float* quats, int64_t n_orient,
                                          const double* s, int64_t P, float recip_extent, const float* c,
                                          void* stream) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (n_orient < 0 || P < 0 || (n_orient > 0 && P > 0 && (!quats || !s))) {
        set_error("setpts_rotated: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (c && p->type != 1) {
        set_error("setpts_rotated: strengths are bound to type-1 plans only");
        return CMTIP_ERR_ARG;
    }
    if (n_orient > 0x7fffffffLL || P > 0x7fffffffLL || n_orient * P > 0x7fffffffLL) {
        set_error("setpts_rotated: at most 2^31 - 1 points per plan");
        return CMTIP_ERR_ARG;
    }
    const int64_t npts = n_orient * P;
    // until the records are built the plan holds no points an execute may use
    p->x = p->y = p->z = nullptr;
    p->npts = npts;
    p->bound = 0;
    p->rot_records = 0;
    if (npts == 0) {
        p->amax_valid = 2;
        p->bound = c ? 1 : 0;
        p->rot_records = 1;
        return CMTIP_OK;
    }
    hipStream_t st = as_stream(stream);
    CMTIP_TRY(ensure_rinv(p, quats, n_orient, st));
    // point j = o * P + pixel: the coordinates of cmtip_rotate_coords, generated where they
    // are binned and copied into the records
    const PtsRotated rot{p->rinv.as<const double>(), s, nullptr, P, inv_ext(recip_extent)};
    CMTIP_TRY(bin_points(p, st, &rot, c, true));
    p->rot_records = 1;
    return CMTIP_OK;
}

extern "C" int cmtip_nufft_points_max_abs(cmtip_nufft_plan p, float* max_abs, void* stream) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (!max_abs) {
        set_error("points_max_abs: null output");
        return CMTIP_ERR_ARG;
    }
    if (p->amax_valid == 2) {
        *max_abs = 0.0f;
        return CMTIP_OK;
    }
    if (!p->amax_valid) {
        set_error("points_max_abs: no point arrays registered since the last setpts");
        return CMTIP_ERR_STATE;
    }
    hipStream_t st = as_stream(stream);
    unsigned bits = 0;
    CMTIP_HIP(hipMemcpyAsync(&bits, p->amax.p, sizeof(bits), hipMemcpyDeviceToHost, st));
    CMTIP_HIP(hipStreamSynchronize(st));
    union {
        unsigned u;
        float f;
    } cv;
    cv.u = bits;
    *max_abs = cv.f;
    return CMTIP_OK;
}

extern "C" int cmtip_nufft_execute(cmtip_nufft_plan p, void* c, void* f, int flags, const float* support,
                                   float scale, void* stream) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    hipStream_t st = as_stream(stream);
    // the adjoint runs the other direction with the opposite sign on the same bins; the
    // sign is flipped on the plan for the duration of the call
    const bool adjoint = (flags & CMTIP_EXEC_ADJOINT) != 0;
    const int type = adjoint ? 3 - p->type : p->type;
    struct SignScope {
        cmtip_nufft_plan_s* p;
        bool flip;
        ~SignScope() {
            if (flip) p->iflag = -p->iflag;
        }
    } sign_scope{p, adjoint};
    if (adjoint) p->iflag = -p->iflag;
    const bool grid_real = flags & CMTIP_GRID_REAL;
    const bool pts_ones = (flags & CMTIP_PTS_ONES) && type == 1;
    const bool pts_imag = (flags & CMTIP_PTS_IMAG) && type == 2;
    const bool pts_bound = (flags & CMTIP_PTS_BOUND) != 0;
    const bool pts_real = (flags & CMTIP_PTS_REAL) || pts_ones || pts_imag || pts_bound;
    if (pts_bound && (type != 1 || pts_ones || !p->bound)) {
        set_error("execute: CMTIP_PTS_BOUND needs a type-1 execute after setpts_strengths");
        return CMTIP_ERR_STATE;
    }
    if (p->bound && !pts_bound && !pts_ones) {
        set_error("execute: the registered records carry strengths (setpts_strengths); call setpts again");
        return CMTIP_ERR_STATE;
    }
    if (pts_ones || pts_bound) c = nullptr;
    if (!f || (p->npts > 0 && !c && !pts_ones && !pts_bound)) {
        set_error("execute: null data pointer");
        return CMTIP_ERR_ARG;
    }
    if (p->npts > 0 && !p->x && !p->rot_records) {
        set_error("execute: setpts not called");
        return CMTIP_ERR_STATE;
    }
    if (type == 2) {
        p->slice_ready = 0;
        // the interpolation weights are real: Re / Im of c interpolate Re / Im of the fine
        // grid, so every form is one or two real passes (complex c: real part into the
        // even floats, then the imaginary part into the odd ones)
        if (pts_real) {
            CMTIP_TRY(grid_to_fine(p, f, grid_real, support, true, st, pts_imag));
            return dispatch_w<InterpLaunch>(p->w, p, (float*)c, 1, (const PtsRotated*)nullptr, st);
        }
        for (int part = 0; part < 2; ++part) {
            CMTIP_TRY(grid_to_fine(p, f, grid_real, support, true, st, part == 1));
            CMTIP_TRY(dispatch_w<InterpLaunch>(p->w, p, (float*)c + part, 2, (const PtsRotated*)nullptr, st));
        }
        return CMTIP_OK;
    }
    // type 1
    if (adjoint && p->npts > 0 && subgrid_lds(p->bs, p->w, true, !pts_real) + kSpreadStaticLds > kMaxLds) {
        set_error("execute: bin size %d of this type-2 plan is too large for the adjoint spread", p->bs);
        return CMTIP_ERR_ARG;
    }
    p->slice_ready = 0;
    const int64_t n3 = (int64_t)p->n * p->n * p->n;
    if (pts_real) {
        // real spreads accumulate in the 64-bit fixed-point grid (converted into fine)
        CMTIP_TRY(p->fine64.ensure(sizeof(unsigned long long) * (size_t)n3));
        CMTIP_TRY(fill_zero(p->fine64.p, 2 * n3, st));
    } else {
        CMTIP_TRY(fill_zero(p->fine.p, 2 * n3, st));
    }
    CMTIP_TRY(dispatch_w<SpreadLaunch>(p->w, p, !pts_real, (const void*)c, pts_bound, st));
    return fine_to_grid(p, pts_real, f, grid_real, support, scale, st);
}

namespace {

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    int init() {
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            set_error("autotune: hipEventCreate failed");
            return CMTIP_ERR_HIP;
        }
        return CMTIP_OK;
    }
    int elapsed(float* ms) {
        if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(ms, a, b) != hipSuccess) {
            set_error("autotune: event timing failed");
            return CMTIP_ERR_HIP;
        }
        return CMTIP_OK;
    }
};

}  // namespace

extern "C" int cmtip_nufft_autotune(cmtip_nufft_plan p, void* c, void* f, int flags, const float* support,
                                    float scale, const int* cands, int n_cand, int reps, float setpts_weight,
                                    float* times_ms, int* best, void* stream) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (!cands || n_cand < 1 || reps < 1 || !(setpts_weight >= 0.0f)) {
        set_error("autotune: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (p->rot_records) {
        set_error("autotune: the points of setpts_rotated are not kept as arrays to bin again; use setpts");
        return CMTIP_ERR_STATE;
    }
    if (p->npts > 0 && !p->x) {
        set_error("autotune: setpts not called");
        return CMTIP_ERR_STATE;
    }
    hipStream_t st = as_stream(stream);
    const int keep_bs = p->bs_req, keep_chunk = p->chunk_req, keep_wpc = p->wpc_req;
    EventPair ev;
    CMTIP_TRY(ev.init());
    int best_i = -1;
    float best_cost = 0.f;
    for (int i = 0; i < n_cand; ++i) {
        float t_set = -1.f, t_exec = -1.f;
        if (cmtip_nufft_set_bins(p, cands[3 * i], cands[3 * i + 1], cands[3 * i + 2]) == CMTIP_OK) {
            if (hipEventRecord(ev.a, st) != hipSuccess) return CMTIP_ERR_HIP;
            CMTIP_TRY(cmtip_nufft_setpts(p, p->npts, p->x, p->y, p->z, st));
            if (hipEventRecord(ev.b, st) != hipSuccess) return CMTIP_ERR_HIP;
            CMTIP_TRY(ev.elapsed(&t_set));
            CMTIP_TRY(cmtip_nufft_execute(p, c, f, flags, support, scale, st));  // warm-up
            if (hipEventRecord(ev.a, st) != hipSuccess) return CMTIP_ERR_HIP;
            for (int r = 0; r < reps; ++r) CMTIP_TRY(cmtip_nufft_execute(p, c, f, flags, support, scale, st));
            if (hipEventRecord(ev.b, st) != hipSuccess) return CMTIP_ERR_HIP;
            CMTIP_TRY(ev.elapsed(&t_exec));
            t_exec /= (float)reps;
            const float cost = setpts_weight * t_set + t_exec;
            if (best_i < 0 || cost < best_cost) {
                best_i = i;
                best_cost = cost;
            }
        }
        if (times_ms) {
            times_ms[2 * i] = t_set;
            times_ms[2 * i + 1] = t_exec;
        }
    }
    if (best_i < 0) {
        // no candidate fits: restore the previous settings and binning
        p->bs_req = keep_bs;
        p->chunk_req = keep_chunk;
        p->wpc_req = keep_wpc;
        CMTIP_TRY(cmtip_nufft_setpts(p, p->npts, p->x, p->y, p->z, st));
        set_error("autotune: no candidate fits this plan");
        return CMTIP_ERR_ARG;
    }
    CMTIP_TRY(cmtip_nufft_set_bins(p, cands[3 * best_i], cands[3 * best_i + 1], cands[3 * best_i + 2]));
    CMTIP_TRY(cmtip_nufft_setpts(p, p->npts, p->x, p->y, p->z, st));
    if (best) *best = best_i;
    return CMTIP_OK;
}

extern "C" int cmtip_nufft_destroy(cmtip_nufft_plan p) {
    if (!p) return CMTIP_OK;
    for (hipEvent_t e : p->tev)
        if (e) (void)hipEventDestroy(e);
    fft1d_free(p->fft);
    p->deconv.release();
    p->fine.release();
    p->tmpA.release();
    p->tmpB.release();
    p->rinv.release();
    p->part.release();
    for (DevBuf* b : {&p->bin_of, &p->count, &p->offset, &p->nchunk, &p->perm, &p->sorted, &p->chunks,
                      &p->ctrl, &p->probe, &p->amax})
        b->release();
    delete p;
    return CMTIP_OK;
}

extern "C" int cmtip_rotate_coords(const float* quats, int64_t n_orient, const double* s, int64_t P,
                                   float recip_extent, float* x, float* y, float* z, void* stream) {
    if (n_orient < 0 || P < 0 || (n_orient > 0 && P > 0 && (!quats || !s || !x || !y || !z))) {
        set_error("rotate_coords: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (n_orient == 0 || P == 0) return CMTIP_OK;
    hipStream_t st = as_stream(stream);
    DevBuf rinv;
    CMTIP_TRY(rinv.ensure((size_t)n_orient * 9 * sizeof(double)));
    hipLaunchKernelGGL(rotinv_kernel, dim3(blocks_for(n_orient)), dim3(kBlock), 0, st, quats, n_orient,
                       rinv.as<double>());
    hipError_t e = hipGetLastError();
    const int64_t total = n_orient * P;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(rotate_coords_kernel, dim3(blocks_for(total)), dim3(kBlock), 0, st,
                           rinv.as<const double>(), s, P, total, inv_ext(recip_extent), x, y, z);
        e = hipGetLastError();
    }
    // the scratch matrix buffer must outlive the kernels
    hipError_t e2 = hipStreamSynchronize(st);
    rinv.release();
    if (e != hipSuccess || e2 != hipSuccess) {
        set_error("rotate_coords: %s", hipGetErrorString(e != hipSuccess ? e : e2));
        return CMTIP_ERR_LAUNCH;
    }
    return CMTIP_OK;
}

extern "C" int cmtip_slice_prepare(cmtip_nufft_plan p, const float* ac, const float* support, void* stream) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (p->type != 2 || !ac) {
        set_error("slice_prepare: needs a type-2 plan and an ac grid");
        return CMTIP_ERR_ARG;
    }
    CMTIP_TRY(grid_to_fine(p, ac, true, support, true, as_stream(stream)));
    p->slice_ready = 1;
    return CMTIP_OK;
}

namespace {

constexpr int kProbeOrients = 16;           // orientations sliced by the route probe
constexpr double kTileMinFit = 0.95;        // tile route when this share of tiles fits its margin,
constexpr int64_t kTileMinPerBin = 64;      // the call has this many tiles per bin
constexpr int kTileMinGrid = 128;           // and the fine grid is at least this wide
constexpr double kFusedMinFit = 0.8;        // else fused tiles for small calls when this share of boxes fits LDS
constexpr int64_t kFusedMaxPts = 1 << 20;   // (small: no sort to amortise); else binned points
constexpr int64_t kBinnedBatchPts = 1 << 27;  // points per binned-route batch (8 B of sort state each)

// Probe for the automatic route on the first orientations (their slices are written
// to out): the share of tiles whose fused box fits LDS and the share within the tile
// route's footprint margin.  One host synchronisation.
int probe_route(cmtip_nufft_plan_s* p, const double* s, const int* order, int64_t P, int64_t n_orient, float ext,
                float* out, hipStream_t st) {
    CMTIP_TRY(p->probe.ensure(3 * sizeof(int)));
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(kBlock), 0, st, p->probe.as<int>(), (int64_t)3, 0);
    CMTIP_LAUNCH_CHECK();
    CMTIP_TRY(dispatch_w<SliceLaunch>(p->w, (const cmtip_nufft_plan_s*)p, s, order, P,
                                      std::min<int64_t>(n_orient, kProbeOrients), ext, out, (uint16_t*)nullptr,
                                      (double*)nullptr, p->probe.as<int>(), st));
    int h[3] = {0, 0, 0};
    if (hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(h, p->probe.p, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("slice_real: route probe failed");
        return CMTIP_ERR_HIP;
    }
    const double tiles = h[1] > 0 ? (double)h[1] : 1.0;
    p->fit_box = h[0] / tiles;
    p->fit_tile = h[2] / tiles;
    return CMTIP_OK;
}

// Automatic route for a call of n_orient orientations from the probed shares
// (measured on 1000 references at M = 41 with 64^2 pixels, 10000 at M = 81 with 128^2
// and 10000 at M = 151: binned tiles win at M = 81, 128^2 pixels; binned points at
// M = 41 and 151; fused tiles for calls of a few orientations).
int auto_route(const cmtip_nufft_plan_s* p, int64_t P, int64_t n_orient) {
    const int64_t nbx = (p->n + kTileBs - 1) / kTileBs;
    const bool tile_reuse = n_orient * ((P + 63) / 64) >= kTileMinPerBin * nbx * nbx * nbx;
    if (p->fit_tile >= kTileMinFit && tile_reuse && p->n >= kTileMinGrid) return 3;
    if (p->fit_box >= kFusedMinFit && n_orient * P <= kFusedMaxPts) return 1;
    return 2;
}

// Kernel timer: event `which` (0 start, 1 end) of the current timed call.
int timer_mark(cmtip_nufft_plan_s* p, int which, hipStream_t st) {
    if (hipEventRecord(p->tev[2 * p->timed + which], st) != hipSuccess) {
        set_error("slice timer: hipEventRecord failed");
        return CMTIP_ERR_HIP;
    }
    if (which == 1) ++p->timed;
    return CMTIP_OK;
}

// Tile-binned route over all orientations (tiles are counted in 32-bit indices).
int slice_tiles(cmtip_nufft_plan_s* p, const double* s, const int* order, int64_t P, int64_t n_orient, float ext,
                float* out, uint16_t* out_bf16, double* part, hipStream_t st, bool timed) {
    const int64_t tpo = (P + 63) / 64, ntiles = tpo * n_orient;
    if (ntiles > 0x7fffffffLL) {
        set_error("slice_real: %lld tiles is too many for the tile route", (long long)ntiles);
        return CMTIP_ERR_ARG;
    }
    const TileRot rot{p->rinv.as<const double>(), s, order, P, tpo, inv_ext(ext)};
    p->x = p->y = p->z = nullptr;  // the bins no longer hold registered points
    p->rot_records = 0;
    p->npts = std::max<int64_t>(p->npts, 1);
    return dispatch_w<TileLaunch>(p->w, p, rot, ntiles, out, out_bf16, part, timed, st);
}

// Binned route: bin sort of the rotated pixels -> LDS-subgrid interpolation, in
// batches of orientations.  Coordinates are regenerated from (rinv, s, order) by the
// count and interpolation kernels (never stored); points are counted in the tile
// order, so the pixels of a wave share few bins.  The plan's registered points are
// dropped.
int slice_binned(cmtip_nufft_plan_s* p, const double* s, const int* order, int64_t P, int64_t n_orient, float ext,
                 float* out, hipStream_t st) {
    const int64_t per = std::max<int64_t>(1, kBinnedBatchPts / P);
    if (std::min(per, n_orient) * P > 0x7fffffffLL) {
        set_error("slice_real: %lld pixels per orientation is too many for the binned route", (long long)P);
        return CMTIP_ERR_ARG;
    }
    p->x = p->y = p->z = nullptr;
    p->rot_records = 0;
    for (int64_t o0 = 0; o0 < n_orient; o0 += per) {
        const int64_t no = std::min(per, n_orient - o0), N = no * P;
        const PtsRotated rot{p->rinv.as<const double>() + 9 * o0, s, order, P, inv_ext(ext)};
        p->npts = N;
        CMTIP_TRY(bin_points(p, st, &rot));
        CMTIP_TRY(dispatch_w<InterpLaunch>(p->w, p, out + o0 * P, 1, &rot, st));
    }
    // npts > 0 with no coordinates: execute reports that setpts is needed
    return CMTIP_OK;
}

}  // namespace

extern "C" int cmtip_slice_set_route(cmtip_nufft_plan p, int route) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (route < 0 || route > 3) {
        set_error("slice_set_route: route %d not in {0 auto, 1 fused, 2 binned, 3 tiles}", route);
        return CMTIP_ERR_ARG;
    }
    p->route_req = route;
    p->route_probed = 0;  // the next automatic call probes again
    return CMTIP_OK;
}

extern "C" int cmtip_slice_timer(cmtip_nufft_plan p, int start, float* total_ms, int* count) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (start) {
        for (hipEvent_t& e : p->tev)
            if (!e && hipEventCreate(&e) != hipSuccess) {
                set_error("slice timer: hipEventCreate failed");
                return CMTIP_ERR_HIP;
            }
        p->timing = 1;
        p->timed = 0;
        return CMTIP_OK;
    }
    float sum = 0.f;
    for (int i = 0; i < p->timed; ++i) {
        float ms = 0.f;
        if (hipEventSynchronize(p->tev[2 * i + 1]) != hipSuccess ||
            hipEventElapsedTime(&ms, p->tev[2 * i], p->tev[2 * i + 1]) != hipSuccess) {
            set_error("slice timer: event timing failed");
            return CMTIP_ERR_HIP;
        }
        sum += ms;
    }
    if (total_ms) *total_ms = sum;
    if (count) *count = p->timed;
    p->timing = 0;
    p->timed = 0;
    return CMTIP_OK;
}

extern "C" int cmtip_slice_get_route(cmtip_nufft_plan p, int* requested, int* last_used) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (requested) *requested = p->route_req;
    if (last_used) *last_used = p->route_used;
    return CMTIP_OK;
}

extern "C" int cmtip_slice_real(cmtip_nufft_plan p, const float* quats, int64_t n_orient, const double* s,
                                const int* order, int64_t P, float recip_extent, float* out, uint16_t* out_bf16,
                                double* moments, void* stream) {
    if (!valid_plan(p)) return CMTIP_ERR_ARG;
    if (!p->slice_ready) {
        set_error("slice_real: cmtip_slice_prepare not called since the last execute");
        return CMTIP_ERR_STATE;
    }
    if (n_orient < 0 || P < 0 || (n_orient > 0 && P > 0 && (!quats || !s || !out))) {
        set_error("slice_real: bad arguments");
        return CMTIP_ERR_ARG;
    }
    if (n_orient == 0 || P == 0) return CMTIP_OK;
    hipStream_t st = as_stream(stream);
    CMTIP_TRY(ensure_rinv(p, quats, n_orient, st));
    int route = p->route_req;
    if (route == 0) {
        if (!(p->route_probed && p->route_s == s && p->route_P == P && p->route_ext == recip_extent &&
              p->route_order == order)) {
            CMTIP_TRY(probe_route(p, s, order, P, n_orient, recip_extent, out, st));
            p->route_probed = 1;
            p->route_s = s;
            p->route_P = P;
            p->route_ext = recip_extent;
            p->route_order = order;
        }
        route = auto_route(p, P, n_orient);
    }
    p->route_used = route;
    // moment partials: one slot per tile of 64 pixels (fused route), per row of 16 lanes
    // of each tile (tile route) or per wave of the finishing pass (binned points); every
    // slot is written by the launch
    double* part = nullptr;
    const int64_t total = n_orient * P;
    const int64_t finish_grid = std::min<int64_t>(4096, (total + kBlock - 1) / kBlock);
    const int64_t ntiles = ((P + 63) / 64) * n_orient;
    const int64_t nslots = route == 2 ? finish_grid * (kBlock / 64) : route == 3 ? 4 * ntiles : ntiles;
    if (moments) {
        CMTIP_TRY(p->part.ensure(sizeof(double) * 2 * (size_t)(nslots + kMomentBlocks)));
        part = p->part.as<double>();
    }
    const bool timed = p->timing && p->timed < 64;
    if (route == 1) {
        if (timed) CMTIP_TRY(timer_mark(p, 0, st));
        CMTIP_TRY(dispatch_w<SliceLaunch>(p->w, (const cmtip_nufft_plan_s*)p, s, order, P, n_orient, recip_extent,
                                          out, out_bf16, part, (int*)nullptr, st));
        if (timed) CMTIP_TRY(timer_mark(p, 1, st));
    } else if (route == 3) {
        CMTIP_TRY(slice_tiles(p, s, order, P, n_orient, recip_extent, out, out_bf16, part, st, timed));
    } else {
        if (timed) CMTIP_TRY(timer_mark(p, 0, st));
        CMTIP_TRY(slice_binned(p, s, order, P, n_orient, recip_extent, out, st));
        if (timed) CMTIP_TRY(timer_mark(p, 1, st));
        if (out_bf16 || part) {
            hipLaunchKernelGGL(slice_finish_kernel, dim3((unsigned)finish_grid), dim3(kBlock), 0, st,
                               (const float*)out, total, out_bf16, part);
            CMTIP_LAUNCH_CHECK();
        }
    }
    if (moments) {
        double* stage = part + 2 * nslots;
        hipLaunchKernelGGL(moments_stage1_kernel, dim3(kMomentBlocks), dim3(kBlock), 0, st, (const double*)part,
                           nslots, stage);
        hipLaunchKernelGGL(moments_stage2_kernel, dim3(1), dim3(64), 0, st, (const double*)stage, moments);
        CMTIP_LAUNCH_CHECK();
    }
    return CMTIP_OK;
}
