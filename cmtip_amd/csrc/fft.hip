// Mixed-radix Stockham FFT axis passes (see fft.h for the contract).
//
// Stage formulation (Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors", SC 2008): with Ns the product of the
// radices already applied, butterfly j of a radix-R stage reads x[j + r*n/R],
// twiddles by exp(sign*2*pi*i*r*(j mod Ns)/(Ns*R)), does an R-point DFT and
// writes to (j div Ns)*Ns*R + (j mod Ns) + r*Ns.  Radices 4, 9, 2, 3 and 5 have
// hand-written butterflies (9 as 3 x 3 in registers, so 162 = 9 * 9 * 2 and 81 = 9 * 9
// take three and two stages); any other prime factor (e.g. 41 or 151, the
// autocorrelation edge lengths whose Fourier-corner projection needs an exact
// M-point DFT) runs a direct O(R^2) DFT stage.
#include <math.h>
#include <stdlib.h>

#include <vector>

#include "fft.h"

namespace cmtip {

namespace {

constexpr int kThreads = 256;
// LDS per workgroup for the two line buffers.  More lines per workgroup make longer
// contiguous runs on strided axes, fewer let more workgroups stay resident: a pass is
// one round of resident workgroups when it fits, and its time is then one
// workgroup's latency.  32 KB (8 lines of 162, 16 of 81; 64-byte runs, 7 workgroups
// per CU) measured best for the autocorrelation operator at M = 81 and equal to 64 KB
// for the NUFFT grids; CMTIP_FFT_LDS_KB overrides it for tuning.
int lds_budget() {
    static int b = 0;
    if (!b) {
        const char* e = getenv("CMTIP_FFT_LDS_KB");
        const int kb = e ? atoi(e) : 32;
        b = (kb >= 16 && kb <= 128 ? kb : 32) * 1024;
    }
    return b;
}

struct PassArgs {
    AxisDesc d;
    int n, nst, lb, lbshift, ldl, ax_contig, ntw;
    uint32_t nlines;
    FastDiv fn, fnv;
    int R[kFftMaxStages], Ns[kFftMaxStages], twoff[kFftMaxStages], omoff[kFftMaxStages];
    FastDiv fnb[kFftMaxStages], fNs[kFftMaxStages];
    const float2* tw;
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// sign * i * a
__device__ __forceinline__ float2 cmuli(float2 a, float sign) { return make_float2(-sign * a.y, sign * a.x); }

template <int R>
__device__ __forceinline__ void butterfly(float2* v, float sg) {
    if constexpr (R == 2) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const float c = -0.5f, s = 0.86602540378443864676f;
        float2 a = cadd(v[1], v[2]), b = csub(v[1], v[2]);
        float2 m = make_float2(v[0].x + c * a.x, v[0].y + c * a.y);
        float2 t = cmuli(cscale(b, s), sg);
        v[0] = cadd(v[0], a);
        v[1] = cadd(m, t);
        v[2] = csub(m, t);
    } else if constexpr (R == 4) {
        float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        float2 t2 = cadd(v[1], v[3]), t3 = cmuli(csub(v[1], v[3]), sg);
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    } else if constexpr (R == 5) {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        float2 m1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
        float2 m2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
        float2 t1 = cmuli(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y), sg);
        float2 t2 = cmuli(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y), sg);
        v[0] = make_float2(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
        v[1] = cadd(m1, t1);
        v[4] = csub(m1, t1);
        v[2] = cadd(m2, t2);
        v[3] = csub(m2, t2);
    } else if constexpr (R == 9) {
        // 3 x 3: n = 3 n1 + n2, k = k1 + 3 k2; 3-point DFTs over n1, twiddles
        // w9^(n2 k1), 3-point DFTs over n2
        float2 a0[3] = {v[0], v[3], v[6]}, a1[3] = {v[1], v[4], v[7]}, a2[3] = {v[2], v[5], v[8]};
        butterfly<3>(a0, sg);
        butterfly<3>(a1, sg);
        butterfly<3>(a2, sg);
        const float c1 = 0.76604444311897803520f, s1 = 0.64278760968653932632f;   // 2 pi / 9
        const float c2 = 0.17364817766693034885f, s2 = 0.98480775301220805936f;   // 4 pi / 9
        const float c4 = -0.93969262078590838405f, s4 = 0.34202014332566873304f;  // 8 pi / 9
        a1[1] = cmul(a1[1], make_float2(c1, sg * s1));
        a1[2] = cmul(a1[2], make_float2(c2, sg * s2));
        a2[1] = cmul(a2[1], make_float2(c2, sg * s2));
        a2[2] = cmul(a2[2], make_float2(c4, sg * s4));
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) {
            float2 b[3] = {a0[k1], a1[k1], a2[k1]};
            butterfly<3>(b, sg);
            v[k1] = b[0];
            v[k1 + 3] = b[1];
            v[k1 + 6] = b[2];
        }
    }
}

constexpr bool has_butterfly(int R) { return R == 2 || R == 3 || R == 4 || R == 5 || R == 9; }

template <int R>
__device__ __forceinline__ void radix_stage(const PassArgs& a, const float2* __restrict__ twt, int s,
                                            const float2* __restrict__ src, float2* __restrict__ dst, float sg) {
    const uint32_t nb = (uint32_t)(a.n / R);
    const uint32_t Ns = (uint32_t)a.Ns[s];
    const uint32_t total = (uint32_t)a.lb * nb;
    const float2* tw = twt + a.twoff[s];
    for (uint32_t b = threadIdx.x; b < total; b += kThreads) {
        const uint32_t line = fdiv(b, a.fnb[s]);
        const uint32_t j = b - line * nb;
        const uint32_t jm = j - fdiv(j, a.fNs[s]) * Ns;
        const float2* sl = src + line * a.ldl;
        float2* dl = dst + line * a.ldl;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = sl[j + r * nb];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tw[jm * (R - 1) + (r - 1)];
                w.y *= -sg;  // table holds exp(-i...); sg = -1 keeps it, +1 conjugates
                v[r] = cmul(v[r], w);
            }
        }
        butterfly<R>(v, sg);
        const uint32_t o = (j - jm) * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) dl[o + r * Ns] = v[r];
    }
}

__device__ void generic_stage(const PassArgs& a, const float2* __restrict__ twt, int s,
                              const float2* __restrict__ src, float2* __restrict__ dst, float sg) {
    const int R = a.R[s];
    const uint32_t nb = (uint32_t)(a.n / R);
    const uint32_t Ns = (uint32_t)a.Ns[s];
    const uint32_t n = (uint32_t)a.n;
    const uint32_t total = (uint32_t)a.lb * n;
    const float2* tw = twt + a.twoff[s];
    const float2* om = twt + a.omoff[s];
    for (uint32_t b = threadIdx.x; b < total; b += kThreads) {
        const uint32_t line = fdiv(b, a.fn);
        const uint32_t rem = b - line * n;
        const uint32_t r = fdiv(rem, a.fnb[s]);
        const uint32_t j = rem - r * nb;
        const uint32_t jm = j - fdiv(j, a.fNs[s]) * Ns;
        const float2* sl = src + line * a.ldl;
        float2 acc = make_float2(0.f, 0.f);
        int e = 0;
        for (int q = 0; q < R; ++q) {
            float2 x = sl[j + q * nb];
            if (q > 0 && Ns > 1) {
                float2 w = tw[jm * (R - 1) + (q - 1)];
                w.y *= -sg;
                x = cmul(x, w);
            }
            float2 w = om[e];
            w.y *= -sg;
            acc = cadd(acc, cmul(x, w));
            e += (int)r;
            if (e >= R) e -= R;
        }
        dst[line * a.ldl + (j - jm) * R + jm + r * Ns] = acc;
    }
}

// Runs all stages (twiddles read from twt); returns the buffer holding the result.
__device__ float2* run_stages(const PassArgs& a, const float2* twt, float2* b0, float2* b1, float sg) {
    float2* src = b0;
    float2* dst = b1;
    // kept rolled: unrolled over kFftMaxStages with every radix inlined in each copy
    // the kernel outgrows the instruction cache
#pragma unroll 1
    for (int s = 0; s < a.nst; ++s) {
        switch (a.R[s]) {
            case 2: radix_stage<2>(a, twt, s, src, dst, sg); break;
            case 3: radix_stage<3>(a, twt, s, src, dst, sg); break;
            case 4: radix_stage<4>(a, twt, s, src, dst, sg); break;
            case 5: radix_stage<5>(a, twt, s, src, dst, sg); break;
            case 9: radix_stage<9>(a, twt, s, src, dst, sg); break;
            default: generic_stage(a, twt, s, src, dst, sg); break;
        }
        __syncthreads();
        float2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

__device__ __forceinline__ void split_elem(const PassArgs& a, uint32_t e, uint32_t& line, uint32_t& p) {
    if (a.ax_contig) {
        line = fdiv(e, a.fn);
        p = e - line * (uint32_t)a.n;
    } else {
        line = e & (uint32_t)(a.lb - 1);
        p = e >> a.lbshift;
    }
}

// Elements per thread per step of the load and store phases: every global load of a
// step is issued before any of its results is used, so a step waits for one memory
// latency instead of kU.
constexpr int kU = 4;
// Mid-pass multipliers a thread prefetches into registers before the forward stages
// (covers lb * n <= kMidPre * kThreads, which the default LDS budget guarantees).
constexpr int kMidPre = 16;

template <bool IN_REAL, int OUT_MODE, bool MID>
__global__ void __launch_bounds__(kThreads) axis_pass_kernel(PassArgs a) {
    if (a.d.skip && *a.d.skip) return;
    extern __shared__ float2 lds[];
    float2* b0 = lds;
    float2* b1 = lds + a.lb * a.ldl;
    float2* twl = lds + 2 * a.lb * a.ldl;
    const AxisDesc& d = a.d;
    const uint32_t n = (uint32_t)a.n;
    const uint32_t line0 = blockIdx.x * (uint32_t)a.lb;
    const uint32_t tot = (uint32_t)a.lb * n;

    // twiddle table to LDS (read by every butterfly) when it fits beside the lines
    for (int i = threadIdx.x; i < a.ntw; i += kThreads) twl[i] = a.tw[i];
    const float2* twt = a.ntw ? twl : a.tw;

    // CG direction update folded into the load: beta and the first-iteration flag
    float cg_beta = 0.f;
    bool cg_first = false;
    if (IN_REAL && d.cg_p) {
        cg_first = d.cg_ctrl[2] == 1;
        cg_beta = cg_first ? 0.f : (float)(d.cg_scal[0] / d.cg_scal[1]);
    }

    // ---- load (zero padding, input weights and masks folded in); positions outside the
    // input segments read element 0 and are zeroed, so the loads are unconditional
    for (uint32_t e0 = threadIdx.x; e0 < tot; e0 += kU * kThreads) {
        float2 val[kU];
        uint32_t slot[kU];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const uint32_t e = e0 + k * kThreads;
            uint32_t line, p;
            split_elem(a, e, line, p);
            slot[k] = line * a.ldl + p;
            const uint32_t gl = line0 + line;
            int g = -1;
            bool conj = false;
            if ((int)p < d.ia) {
                g = d.ig1 + (int)p;
            } else if ((int)p >= (int)n - d.ib) {
                conj = d.in_herm != 0;
                g = conj ? (int)n - (int)p : d.ig2 + ((int)p - ((int)n - d.ib));
            }
            const bool ok = e < tot && gl < a.nlines && g >= 0;
            int64_t off = 0;
            if (ok) {
                const uint32_t u = fdiv(gl, a.fnv);
                const uint32_t v = gl - u * (uint32_t)d.nv;
                off = (int64_t)u * d.in_su + (int64_t)v * d.in_sv + (int64_t)g * d.in_sa;
            }
            float2 x;
            if (IN_REAL) {
                float r = reinterpret_cast<const float*>(d.in)[off];
                if (d.cg_p) {
                    const float pv = d.cg_p[off];
                    r = cg_first ? r : cg_beta * pv + r;
                    if (ok) d.cg_p[off] = r;
                }
                x = make_float2(r, 0.f);
            } else {
                x = reinterpret_cast<const float2*>(d.in)[off];
                if (conj) x.y = -x.y;
            }
            float wgt = ok ? 1.f : 0.f;
            if (d.in_w) wgt *= d.in_w[ok ? p : 0];
            if (d.in_mask) wgt *= d.in_mask[off];
            val[k] = cscale(x, wgt);
        }
#pragma unroll
        for (int k = 0; k < kU; ++k)
            if (e0 + k * kThreads < tot) b0[slot[k]] = val[k];
    }

    // mid-pass multipliers for this thread's elements, fetched while the forward stages run
    float2 midv[MID ? kMidPre : 1];
    const bool mid_pre = MID && tot <= (uint32_t)(kMidPre * kThreads);
    if (MID && mid_pre && !d.mid_abs2) {
#pragma unroll
        for (int k = 0; k < kMidPre; ++k) {
            const uint32_t e = threadIdx.x + k * kThreads;
            midv[k] = make_float2(0.f, 0.f);
            if (e < tot) {
                uint32_t line, p;
                split_elem(a, e, line, p);
                const uint32_t gl = line0 + line;
                if (gl < a.nlines) {
                    const uint32_t u = fdiv(gl, a.fnv);
                    const uint32_t v = gl - u * (uint32_t)d.nv;
                    const int64_t off = (int64_t)u * d.mid_su + (int64_t)v * d.mid_sv + (int64_t)p * d.mid_sa;
                    if (d.mid_c)
                        midv[k] = d.mid_c[off];
                    else
                        midv[k].x = d.mid_amp ? d.mid_amp[off] : d.mid_r[off];
                }
            }
        }
    }
    __syncthreads();

    const float sg = (float)d.sign;
    float2* res = run_stages(a, twt, b0, b1, sg);

    if (MID) {
        auto apply = [&](float2 x, float2 m) -> float2 {
            if (d.mid_amp) {
                const float am = m.x;
                if (am >= 0.f) {
                    const float r = sqrtf(x.x * x.x + x.y * x.y);
                    x = r > 0.f ? cscale(x, am / r) : make_float2(am, 0.f);
                }
            } else if (d.mid_abs2) {
                x = make_float2(x.x * x.x + x.y * x.y, 0.f);
            } else if (d.mid_c) {
                x = cmul(x, m);
            } else {
                x = cscale(x, m.x);
            }
            return x;
        };
        if (mid_pre) {
#pragma unroll
            for (int k = 0; k < kMidPre; ++k) {
                const uint32_t e = threadIdx.x + k * kThreads;
                if (e >= tot) break;
                uint32_t line, p;
                split_elem(a, e, line, p);
                if (line0 + line >= a.nlines) continue;
                float2& x = res[line * a.ldl + p];
                x = apply(x, midv[k]);
            }
        } else {
            for (uint32_t e = threadIdx.x; e < tot; e += kThreads) {
                uint32_t line, p;
                split_elem(a, e, line, p);
                const uint32_t gl = line0 + line;
                if (gl >= a.nlines) continue;
                const uint32_t u = fdiv(gl, a.fnv);
                const uint32_t v = gl - u * (uint32_t)d.nv;
                const int64_t off = (int64_t)u * d.mid_su + (int64_t)v * d.mid_sv + (int64_t)p * d.mid_sa;
                float2 m = make_float2(0.f, 0.f);
                if (d.mid_c)
                    m = d.mid_c[off];
                else if (!d.mid_abs2)
                    m.x = d.mid_amp ? d.mid_amp[off] : d.mid_r[off];
                float2& x = res[line * a.ldl + p];
                x = apply(x, m);
            }
        }
        __syncthreads();
        float2* other = (res == b0) ? b1 : b0;
        res = run_stages(a, twt, res, other, -sg);
    }

    // ---- store (cropping, output weights, masks and axpy folded in); the step's side
    // inputs (masks, axpy source, dot source, the array updated in place) are loaded
    // unconditionally from element 0 where the position is not stored
    double dot_acc = 0.0;
    for (uint32_t e0 = threadIdx.x; e0 < tot; e0 += kU * kThreads) {
        int64_t offs[kU];
        bool oks[kU];
        float wg[kU], addv[kU], dotv[kU], oldv[kU], supv[kU];
        float2 xs[kU];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const uint32_t e = e0 + k * kThreads;
            uint32_t line, p;
            split_elem(a, e, line, p);
            const uint32_t gl = line0 + line;
            int g = -1;
            if ((int)p < d.oa)
                g = d.og1 + (int)p;
            else if ((int)p >= (int)n - d.ob)
                g = d.og2 + ((int)p - ((int)n - d.ob));
            const bool ok = e < tot && gl < a.nlines && g >= 0;
            int64_t off = 0;
            if (ok) {
                const uint32_t u = fdiv(gl, a.fnv);
                const uint32_t v = gl - u * (uint32_t)d.nv;
                off = (int64_t)u * d.out_su + (int64_t)v * d.out_sv + (int64_t)g * d.out_sa;
            }
            offs[k] = off;
            oks[k] = ok;
            float w = d.out_scale;
            if (d.out_w) w *= d.out_w[ok ? p : 0];
            if (d.out_mask) w *= d.out_mask[off];
            wg[k] = w;
            xs[k] = ok ? res[line * a.ldl + p] : make_float2(0.f, 0.f);
            if (OUT_MODE == 1 || OUT_MODE == 2) {
                addv[k] = d.add_src ? d.add_src[off] : 0.f;
                dotv[k] = d.dot_part ? d.dot_src[off] : 0.f;
                oldv[k] = (OUT_MODE == 1 && d.upd_mode) ? reinterpret_cast<const float*>(d.out)[off] : 0.f;
                supv[k] = (OUT_MODE == 1 && d.upd_mode) ? d.upd_support[off] : 0.f;
            }
        }
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            if (!oks[k]) continue;
            const int64_t off = offs[k];
            const float2 x = cscale(xs[k], wg[k]);
            if (OUT_MODE == 0) {
                reinterpret_cast<float2*>(d.out)[off] = x;
            } else if (OUT_MODE == 3) {
                reinterpret_cast<float*>(d.out)[off] = sqrtf(x.x * x.x + x.y * x.y);
            } else {
                float y = OUT_MODE == 1 ? x.x : x.y;
                if (d.add_src) y += d.add_scale * addv[k];
                if (OUT_MODE == 1 && d.upd_mode) {
                    const bool keep = supv[k] > 0.f && y > 0.f;
                    float r;
                    if (d.upd_mode == 1) {
                        r = keep ? y : 0.f;
                        if (y > d.upd_rho_max) r = d.upd_rho_max;
                    } else {
                        r = keep ? y : oldv[k] - d.upd_beta * y;
                        if (y > d.upd_rho_max) r += 2.f * d.upd_beta * y - d.upd_rho_max;
                    }
                    y = r;
                }
                reinterpret_cast<float*>(d.out)[off] = y;
                if (d.dot_part) dot_acc += (double)dotv[k] * (double)y;
            }
        }
    }
    if (OUT_MODE != 0 && OUT_MODE != 3 && d.dot_part) {
        // fixed-order workgroup sum: wave shuffles, then the waves in index order
        __shared__ double wsum[kThreads / 64];
        for (int o = 32; o > 0; o >>= 1) dot_acc += __shfl_down(dot_acc, o, 64);
        if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = dot_acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < kThreads / 64; ++w) t += wsum[w];
            d.dot_part[blockIdx.x] = t;
        }
    }
}

// lines per workgroup: a power of two that keeps two LDS line buffers under budget
int lines_per_group(int n) {
    const int ldl = n | 1;
    int lb = 1;
    const int budget = lds_budget();
    while (lb < 64 && 2 * (2 * lb) * ldl * (int)sizeof(float2) <= budget) lb *= 2;
    return lb;
}

void factorize(int n, std::vector<int>& f) {
    f.clear();
    while (n % 4 == 0) { f.push_back(4); n /= 4; }
    while (n % 9 == 0) { f.push_back(9); n /= 9; }
    while (n % 2 == 0) { f.push_back(2); n /= 2; }
    while (n % 3 == 0) { f.push_back(3); n /= 3; }
    while (n % 5 == 0) { f.push_back(5); n /= 5; }
    for (int p = 7; n > 1; p += 2)
        while (n % p == 0) { f.push_back(p); n /= p; }
}

}  // namespace

int fft1d_init(Fft1D& p, int n) {
    if (n < 1 || n > 8192) {
        set_error("fft1d_init: unsupported length %d", n);
        return CMTIP_ERR_ARG;
    }
    std::vector<int> f;
    factorize(n, f);
    if ((int)f.size() > kFftMaxStages) {
        set_error("fft1d_init: too many stages for %d", n);
        return CMTIP_ERR_ARG;
    }
    p.n = n;
    p.nst = (int)f.size();
    std::vector<float2> tw;
    int Ns = 1;
    for (int s = 0; s < p.nst; ++s) {
        const int R = f[s];
        p.R[s] = R;
        p.Ns[s] = Ns;
        p.twoff[s] = (int)tw.size();
        for (int jm = 0; jm < Ns; ++jm)
            for (int r = 1; r < R; ++r) {
                const double ang = -2.0 * M_PI * (double)jm * (double)r / ((double)Ns * R);
                tw.push_back(make_float2((float)cos(ang), (float)sin(ang)));
            }
        p.omoff[s] = (int)tw.size();
        if (!has_butterfly(R))
            for (int m = 0; m < R; ++m) {
                const double ang = -2.0 * M_PI * (double)m / (double)R;
                tw.push_back(make_float2((float)cos(ang), (float)sin(ang)));
            }
        Ns *= R;
    }
    if (tw.empty()) tw.push_back(make_float2(1.f, 0.f));
    p.ntw = (int)tw.size();
    CMTIP_TRY(p.tw.ensure(tw.size() * sizeof(float2)));
    CMTIP_HIP(hipMemcpy(p.tw.p, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    return CMTIP_OK;
}

void fft1d_free(Fft1D& p) { p.tw.release(); }

int64_t fft_axis_grid(const Fft1D& plan, const AxisDesc& d) {
    const int64_t lb = lines_per_group(plan.n);
    return ((int64_t)d.nu * d.nv + lb - 1) / lb;
}

int fft_axis_pass(const Fft1D& plan, const AxisDesc& d, hipStream_t stream) {
    PassArgs a;
    a.d = d;
    a.n = plan.n;
    a.nst = plan.nst;
    const int64_t nlines = (int64_t)d.nu * d.nv;
    if (nlines <= 0) return CMTIP_OK;
    if (d.ia + d.ib > plan.n || d.oa + d.ob > plan.n ||
        ((d.mid_c != nullptr) + (d.mid_r != nullptr) + (d.mid_amp != nullptr) + (d.mid_abs2 != 0)) > 1 ||
        (d.upd_mode && (d.out_mode != 1 || !d.upd_support)) || d.out_mode < 0 || d.out_mode > 3 ||
        (d.in_real && d.out_mode == 3) || (d.cg_p && (!d.in_real || !d.cg_scal || !d.cg_ctrl)) ||
        (d.dot_part && (!d.dot_src || (d.out_mode != 1 && d.out_mode != 2))) || (d.in_herm && d.in_real)) {
        set_error("fft_axis_pass: bad segments (n=%d ia=%d ib=%d oa=%d ob=%d)", plan.n, d.ia, d.ib,
                  d.oa, d.ob);
        return CMTIP_ERR_ARG;
    }
    if (nlines * (int64_t)d.nv >= ((int64_t)1 << 32) || nlines >= ((int64_t)1 << 31)) {
        set_error("fft_axis_pass: too many lines");
        return CMTIP_ERR_ARG;
    }
    const int ldl = plan.n | 1;
    const int lb = lines_per_group(plan.n);
    int lbshift = 0;
    while ((1 << lbshift) < lb) ++lbshift;
    a.lb = lb;
    a.lbshift = lbshift;
    a.ldl = ldl;
    a.ax_contig = (d.in_sa == 1 || d.out_sa == 1) ? 1 : 0;
    a.nlines = (uint32_t)nlines;
    a.fn = make_fastdiv((uint32_t)plan.n);
    a.fnv = make_fastdiv((uint32_t)d.nv);
    for (int s = 0; s < plan.nst; ++s) {
        a.R[s] = plan.R[s];
        a.Ns[s] = plan.Ns[s];
        a.twoff[s] = plan.twoff[s];
        a.omoff[s] = plan.omoff[s];
        a.fnb[s] = make_fastdiv((uint32_t)(plan.n / plan.R[s]));
        a.fNs[s] = make_fastdiv((uint32_t)plan.Ns[s]);
    }
    a.tw = plan.tw.as<const float2>();
    a.ntw = ((size_t)2 * lb * ldl + plan.ntw) * sizeof(float2) <= 156 * 1024 ? plan.ntw : 0;
    const size_t lds = ((size_t)2 * lb * ldl + a.ntw) * sizeof(float2);
    const unsigned grid = (unsigned)((nlines + lb - 1) / lb);
    const bool mid = d.mid_c || d.mid_r || d.mid_amp || d.mid_abs2;
#define CMTIP_AXIS_LAUNCH(IR, OM, MD)                                                                       \
    do {                                                                                                    \
        static size_t attr_bytes_ = 64 * 1024; /* dynamic LDS the kernel accepts */                         \
        if (lds > attr_bytes_) {                                                                            \
            if (hipFuncSetAttribute((const void*)axis_pass_kernel<IR, OM, MD>,                              \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {  \
                (void)hipGetLastError();                                                                    \
                set_error("fft_axis_pass: %zu bytes of LDS per workgroup refused", lds);                    \
                return CMTIP_ERR_LAUNCH;                                                                    \
            }                                                                                               \
            attr_bytes_ = lds;                                                                              \
        }                                                                                                   \
        hipLaunchKernelGGL((axis_pass_kernel<IR, OM, MD>), dim3(grid), dim3(kThreads), lds, stream, a);    \
    } while (0)
    if (d.in_real) {
        if (d.out_mode == 0) {
            if (mid) CMTIP_AXIS_LAUNCH(true, 0, true); else CMTIP_AXIS_LAUNCH(true, 0, false);
        } else if (d.out_mode == 1) {
            if (mid) CMTIP_AXIS_LAUNCH(true, 1, true); else CMTIP_AXIS_LAUNCH(true, 1, false);
        } else {
            if (mid) CMTIP_AXIS_LAUNCH(true, 2, true); else CMTIP_AXIS_LAUNCH(true, 2, false);
        }
    } else {
        if (d.out_mode == 0) {
            if (mid) CMTIP_AXIS_LAUNCH(false, 0, true); else CMTIP_AXIS_LAUNCH(false, 0, false);
        } else if (d.out_mode == 1) {
            if (mid) CMTIP_AXIS_LAUNCH(false, 1, true); else CMTIP_AXIS_LAUNCH(false, 1, false);
        } else if (d.out_mode == 2) {
            if (mid) CMTIP_AXIS_LAUNCH(false, 2, true); else CMTIP_AXIS_LAUNCH(false, 2, false);
        } else {
            if (mid) CMTIP_AXIS_LAUNCH(false, 3, true); else CMTIP_AXIS_LAUNCH(false, 3, false);
        }
    }
#undef CMTIP_AXIS_LAUNCH
    CMTIP_LAUNCH_CHECK();
    return CMTIP_OK;
}

}  // namespace cmtip
