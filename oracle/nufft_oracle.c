/*
 * FP64 CPU oracle for the 3-D type-1 / type-2 NUFFT used by the cmtip hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * What it restates (the reference delegates both transforms to finufft, which
 * is not vendored with the reference; see cmtip/nufft/forward.py:40 and
 * cmtip/nufft/adjoint.py:33):
 *   type 2 (forward):  c_j = sum_k f_k exp(-i k.x_j)          (isign -1)
 *   type 1 (adjoint):  f_k = sum_j c_j exp(+i k.x_j)          (isign +1)
 *   k_d in [-floor(M/2), ceil(M/2)), array index i <-> k = i - floor(M/2)
 *   (finufft's default mode ordering, pinned by cmtip/tests/test_nufft.py:31-49).
 *
 * Two independent routes are provided:
 *   - oracle_nudft2 / oracle_nudft1: exact direct sums (O(N M^3)), the ground
 *     truth for small problems;
 *   - oracle_interp / oracle_spread: the exponential-of-semicircle (ES) kernel
 *     gridding steps of Barnett, Magland & af Klinteberg, SIAM J. Sci. Comput.
 *     41(5) C479 (2019).  The FFT and deconvolution steps are done in numpy
 *     (oracle/nufft.py), so the fast route is checked against the direct sums.
 *
 * Kernel, in fine-grid units z (|z| <= w/2):  phi(z) = exp(beta (sqrt(1-(2z/w)^2) - 1)).
 * Point x (radians) sits at fine-grid coordinate t = x n / (2 pi); it touches
 * the w grid nodes l = ceil(t - w/2) ... ceil(t - w/2) + w - 1 (mod n).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXW 32

static inline double es_phi(double z, double w, double beta) {
    double t = 2.0 * z / w;
    double s = 1.0 - t * t;
    if (s < 0.0) return 0.0;
    return exp(beta * (sqrt(s) - 1.0));
}

static inline int64_t wrap(int64_t l, int64_t n) {
    l %= n;
    return l < 0 ? l + n : l;
}

static void kernel_row(double x, int n, int w, double beta, int64_t *idx, double *wt) {
    double t = x * (double)n / (2.0 * M_PI);
    int64_t l0 = (int64_t)ceil(t - 0.5 * w);
    for (int a = 0; a < w; ++a) {
        idx[a] = wrap(l0 + a, n);
        wt[a] = es_phi((double)(l0 + a) - t, (double)w, beta);
    }
}

/* Interpolate the complex fine grid g (n^3, row-major, last axis fastest) at N points. */
void oracle_interp(int n, int w, double beta, const double *g_re, const double *g_im,
                   int64_t N, const double *x, const double *y, const double *z,
                   double *c_re, double *c_im) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t j = 0; j < N; ++j) {
        int64_t ix[MAXW], iy[MAXW], iz[MAXW];
        double wx[MAXW], wy[MAXW], wz[MAXW];
        kernel_row(x[j], n, w, beta, ix, wx);
        kernel_row(y[j], n, w, beta, iy, wy);
        kernel_row(z[j], n, w, beta, iz, wz);
        double sr = 0.0, si = 0.0;
        for (int a = 0; a < w; ++a) {
            double pr = 0.0, pi = 0.0;
            for (int b = 0; b < w; ++b) {
                const int64_t base = (ix[a] * n + iy[b]) * (int64_t)n;
                double qr = 0.0, qi = 0.0;
                for (int c = 0; c < w; ++c) {
                    qr += wz[c] * g_re[base + iz[c]];
                    qi += wz[c] * g_im[base + iz[c]];
                }
                pr += wy[b] * qr;
                pi += wy[b] * qi;
            }
            sr += wx[a] * pr;
            si += wx[a] * pi;
        }
        c_re[j] = sr;
        c_im[j] = si;
    }
}

/* Spread N complex strengths onto the fine grid b (n^3), accumulating (b is not cleared).
 * Threads spread into private slabs of the x axis: each thread owns a contiguous
 * range of x planes and visits every point, keeping only the planes it owns. */
void oracle_spread(int n, int w, double beta, int64_t N, const double *x, const double *y,
                   const double *z, const double *c_re, const double *c_im,
                   double *b_re, double *b_im) {
#pragma omp parallel
    {
        int nth = 1, tid = 0;
#ifdef _OPENMP
        nth = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        int64_t lo = (int64_t)n * tid / nth, hi = (int64_t)n * (tid + 1) / nth;
        for (int64_t j = 0; j < N; ++j) {
            int64_t ix[MAXW], iy[MAXW], iz[MAXW];
            double wx[MAXW], wy[MAXW], wz[MAXW];
            kernel_row(x[j], n, w, beta, ix, wx);
            int any = 0;
            for (int a = 0; a < w; ++a) any |= (ix[a] >= lo && ix[a] < hi);
            if (!any) continue;
            kernel_row(y[j], n, w, beta, iy, wy);
            kernel_row(z[j], n, w, beta, iz, wz);
            for (int a = 0; a < w; ++a) {
                if (ix[a] < lo || ix[a] >= hi) continue;
                for (int b = 0; b < w; ++b) {
                    const double wab = wx[a] * wy[b];
                    const double vr = wab * c_re[j], vi = wab * c_im[j];
                    const int64_t base = (ix[a] * n + iy[b]) * (int64_t)n;
                    for (int c = 0; c < w; ++c) {
                        b_re[base + iz[c]] += wz[c] * vr;
                        b_im[base + iz[c]] += wz[c] * vi;
                    }
                }
            }
        }
    }
}

/* Exact type-2 sum.  f is (M,M,M) complex, row-major; out c has N entries. */
void oracle_nudft2(int M, const double *f_re, const double *f_im, int64_t N,
                   const double *x, const double *y, const double *z,
                   double *c_re, double *c_im) {
    const int k0 = M / 2;
#pragma omp parallel
    {
        double *ex_r = malloc(sizeof(double) * 3 * M), *ex_i = malloc(sizeof(double) * 3 * M);
#pragma omp for schedule(static)
        for (int64_t j = 0; j < N; ++j) {
            const double xs[3] = {x[j], y[j], z[j]};
            for (int d = 0; d < 3; ++d)
                for (int i = 0; i < M; ++i) {
                    double ph = -(double)(i - k0) * xs[d];
                    ex_r[d * M + i] = cos(ph);
                    ex_i[d * M + i] = sin(ph);
                }
            double sr = 0.0, si = 0.0;
            for (int a = 0; a < M; ++a) {
                double pr = 0.0, pi = 0.0;
                for (int b = 0; b < M; ++b) {
                    const int64_t base = ((int64_t)a * M + b) * M;
                    double qr = 0.0, qi = 0.0;
                    for (int c = 0; c < M; ++c) {
                        double er = ex_r[2 * M + c], ei = ex_i[2 * M + c];
                        qr += f_re[base + c] * er - f_im[base + c] * ei;
                        qi += f_re[base + c] * ei + f_im[base + c] * er;
                    }
                    double er = ex_r[M + b], ei = ex_i[M + b];
                    pr += qr * er - qi * ei;
                    pi += qr * ei + qi * er;
                }
                double er = ex_r[a], ei = ex_i[a];
                sr += pr * er - pi * ei;
                si += pr * ei + pi * er;
            }
            c_re[j] = sr;
            c_im[j] = si;
        }
        free(ex_r);
        free(ex_i);
    }
}

/* Exact type-1 sum onto an (M,M,M) grid (overwritten). */
void oracle_nudft1(int M, int64_t N, const double *x, const double *y, const double *z,
                   const double *c_re, const double *c_im, double *f_re, double *f_im) {
    const int k0 = M / 2;
    /* per-point exponential tables for the second and third axes: E[j*M + i] */
    double *eyr = malloc(sizeof(double) * (size_t)N * M), *eyi = malloc(sizeof(double) * (size_t)N * M);
    double *ezr = malloc(sizeof(double) * (size_t)N * M), *ezi = malloc(sizeof(double) * (size_t)N * M);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            eyr[j * M + i] = cos((double)(i - k0) * y[j]);
            eyi[j * M + i] = sin((double)(i - k0) * y[j]);
            ezr[j * M + i] = cos((double)(i - k0) * z[j]);
            ezi[j * M + i] = sin((double)(i - k0) * z[j]);
        }
#pragma omp parallel
    {
        double *tr = malloc(sizeof(double) * (size_t)N), *ti = malloc(sizeof(double) * (size_t)N);
        double *ur = malloc(sizeof(double) * (size_t)N), *ui = malloc(sizeof(double) * (size_t)N);
#pragma omp for schedule(dynamic, 1)
        for (int a = 0; a < M; ++a) {
            for (int64_t j = 0; j < N; ++j) {
                double ph = (double)(a - k0) * x[j];
                double cr = cos(ph), ci = sin(ph);
                tr[j] = c_re[j] * cr - c_im[j] * ci;
                ti[j] = c_re[j] * ci + c_im[j] * cr;
            }
            for (int b = 0; b < M; ++b) {
                for (int64_t j = 0; j < N; ++j) {
                    double er = eyr[j * M + b], ei = eyi[j * M + b];
                    ur[j] = tr[j] * er - ti[j] * ei;
                    ui[j] = tr[j] * ei + ti[j] * er;
                }
                for (int c = 0; c < M; ++c) {
                    double sr = 0.0, si = 0.0;
                    for (int64_t j = 0; j < N; ++j) {
                        double er = ezr[j * M + c], ei = ezi[j * M + c];
                        sr += ur[j] * er - ui[j] * ei;
                        si += ur[j] * ei + ui[j] * er;
                    }
                    const int64_t o = ((int64_t)a * M + b) * M + c;
                    f_re[o] = sr;
                    f_im[o] = si;
                }
            }
        }
        free(tr); free(ti); free(ur); free(ui);
    }
    free(eyr); free(eyi); free(ezr); free(ezi);
}

/* Exact type-1 sums at a tensor product of selected modes: out[(a*nb + b)*nc + c] =
 * sum_j c_j exp(+i (ka[a] x_j + kb[b] y_j + kc[c] z_j)) for integer mode numbers ka, kb,
 * kc (k = i - floor(M/2) for array index i).  Each point costs na + nb + nc sincos
 * evaluations and na nb nc complex products, which keeps exact checks of a few hundred
 * modes affordable at tens of millions of points (the full-grid sum above is O(N M^3)). */
void oracle_nudft1_modes(int64_t N, const double *x, const double *y, const double *z,
                         const double *c_re, const double *c_im, int na, const int *ka, int nb,
                         const int *kb, int nc, const int *kc, double *out_re, double *out_im) {
    const int nt = na * nb * nc;
    for (int t = 0; t < nt; ++t) out_re[t] = out_im[t] = 0.0;
#pragma omp parallel
    {
        double *ar = calloc((size_t)nt, sizeof(double)), *ai = calloc((size_t)nt, sizeof(double));
        double ex_r[MAXW * 8], ex_i[MAXW * 8], ey_r[MAXW * 8], ey_i[MAXW * 8], ez_r[MAXW * 8], ez_i[MAXW * 8];
#pragma omp for schedule(static)
        for (int64_t j = 0; j < N; ++j) {
            for (int a = 0; a < na; ++a) { ex_r[a] = cos(ka[a] * x[j]); ex_i[a] = sin(ka[a] * x[j]); }
            for (int b = 0; b < nb; ++b) { ey_r[b] = cos(kb[b] * y[j]); ey_i[b] = sin(kb[b] * y[j]); }
            for (int c = 0; c < nc; ++c) { ez_r[c] = cos(kc[c] * z[j]); ez_i[c] = sin(kc[c] * z[j]); }
            const double vr = c_re[j], vi = c_im ? c_im[j] : 0.0;
            for (int a = 0; a < na; ++a) {
                const double pr = vr * ex_r[a] - vi * ex_i[a], pi = vr * ex_i[a] + vi * ex_r[a];
                for (int b = 0; b < nb; ++b) {
                    const double qr = pr * ey_r[b] - pi * ey_i[b], qi = pr * ey_i[b] + pi * ey_r[b];
                    double *dr = ar + ((int64_t)a * nb + b) * nc, *di = ai + ((int64_t)a * nb + b) * nc;
                    for (int c = 0; c < nc; ++c) {
                        dr[c] += qr * ez_r[c] - qi * ez_i[c];
                        di[c] += qr * ez_i[c] + qi * ez_r[c];
                    }
                }
            }
        }
#pragma omp critical
        for (int t = 0; t < nt; ++t) {
            out_re[t] += ar[t];
            out_im[t] += ai[t];
        }
        free(ar);
        free(ai);
    }
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
