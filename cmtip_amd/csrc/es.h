// Exponential-of-semicircle (ES) spreading kernel (Barnett, Magland & af Klinteberg,
// SIAM J. Sci. Comput. 41(5) C479, 2019), upsampling factor 2.
//
// In fine-grid units z (|z| <= w/2):  phi(z) = exp(beta * (sqrt(1 - (2z/w)^2) - 1)).
// A point at fine-grid coordinate t = x * n / (2 pi) touches the w nodes
// l0, ..., l0 + w - 1 with l0 = ceil(t - w/2) (wrapped mod n).
#pragma once
#include <math.h>

#include <algorithm>
#include <vector>

namespace cmtip {

struct EsParams {
    int w;
    double beta;
};

// Width rule of the paper for upsampling 2: w = ceil(log10(1/eps)) + 1, beta = 2.30 w.
inline EsParams es_params(double eps) {
    if (!(eps > 0)) eps = 1e-6;
    int w = (int)ceil(log10(1.0 / eps)) + 1;
    if (w < 2) w = 2;
    if (w > 16) w = 16;
    return EsParams{w, 2.30 * w};
}

inline double es_phi_host(double z, int w, double beta) {
    const double t = 2.0 * z / w;
    const double s = 1.0 - t * t;
    return s < 0.0 ? 0.0 : exp(beta * (sqrt(s) - 1.0));
}

// Polynomial form of the w taps of one axis (w <= kEsPolyMaxW).  With tau = t - l0 in
// (w/2 - 1, w/2] and y = tau - (w - 1)/2 in (-1/2, 1/2], tap a weighs
// f_a(y) = phi(a - tau); the mirror tap satisfies f_{w-1-a}(y) = f_a(-y).  Each f_a
// (a < w/2) is the degree-(w + 1) polynomial interpolating it at the Chebyshev points of
// [-1/2, 1/2]; it is stored as its even part E(y^2) and odd part y O(y^2), so the pair
// is E + y O and E - y O.  Odd w adds the even part of the middle tap.  Error against
// phi with the float32 coefficients (max over taps and y; eps = 10^(1 - w), the
// tolerance a width serves): w = 2 .. 8: 1.4e-2, 3.8e-4, 8.5e-5, 5.4e-6, 6.2e-7, 5.9e-8,
// 2.7e-8, i.e. below eps / 7 for w <= 7 and eps / 3.7 at w = 8 (float32 rounding).
// At w = 7 the outer taps (the square-root edge of phi) carry it; the inner ones are
// below 3.1e-8.
constexpr int kEsPolyMaxW = 8;
constexpr int kEsPolyCoefs = 40;  // w = 8: 4 pairs x 10

inline int es_poly_even(int w) { return (w + 1) / 2 + 1; }  // coefficients of y^0, y^2, ...
inline int es_poly_odd(int w) { return (w + 2) / 2; }        // coefficients of y^1, y^3, ...

// Monomial coefficients (y^0 .. y^D) of the interpolant of tap a at D + 1 Chebyshev points.
inline void es_tap_fit(int w, double beta, int a, int D, double* c) {
    const int n = D + 1;
    std::vector<double> A((size_t)n * (n + 1));
    for (int k = 0; k < n; ++k) {
        const double y = 0.5 * cos((2.0 * k + 1.0) * M_PI / (2.0 * n));
        const double tau = y + 0.5 * (w - 1);
        double pw = 1.0;
        for (int j = 0; j < n; ++j, pw *= y) A[(size_t)k * (n + 1) + j] = pw;
        A[(size_t)k * (n + 1) + n] = es_phi_host(a - tau, w, beta);
    }
    // Gaussian elimination with partial pivoting
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (fabs(A[(size_t)r * (n + 1) + col]) > fabs(A[(size_t)piv * (n + 1) + col])) piv = r;
        for (int j = 0; j <= n; ++j) std::swap(A[(size_t)col * (n + 1) + j], A[(size_t)piv * (n + 1) + j]);
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double f = A[(size_t)r * (n + 1) + col] / A[(size_t)col * (n + 1) + col];
            for (int j = col; j <= n; ++j) A[(size_t)r * (n + 1) + j] -= f * A[(size_t)col * (n + 1) + j];
        }
    }
    for (int j = 0; j < n; ++j) c[j] = A[(size_t)j * (n + 1) + n] / A[(size_t)j * (n + 1) + j];
}

// Fills c (kEsPolyCoefs floats) in the layout described above; false for w > kEsPolyMaxW.
inline bool es_poly(int w, double beta, float* c) {
    for (int i = 0; i < kEsPolyCoefs; ++i) c[i] = 0.0f;
    if (w < 2 || w > kEsPolyMaxW) return false;
    const int D = w + 1, ne = es_poly_even(w), no = es_poly_odd(w);
    double m[kEsPolyMaxW + 2];
    int k = 0;
    for (int a = 0; a < w / 2; ++a) {
        es_tap_fit(w, beta, a, D, m);
        for (int j = 0; j < ne; ++j) c[k++] = (float)(2 * j <= D ? m[2 * j] : 0.0);
        for (int j = 0; j < no; ++j) c[k++] = (float)(2 * j + 1 <= D ? m[2 * j + 1] : 0.0);
    }
    if (w & 1) {
        es_tap_fit(w, beta, w / 2, D, m);
        for (int j = 0; j < ne; ++j) c[k++] = (float)(2 * j <= D ? m[2 * j] : 0.0);
    }
    return true;
}

// The ES kernel as the kernels take it (by value): the exponent scale beta log2(e) of
// the direct form and the tap polynomials of es_poly.
struct EsKernel {
    float beta_l2e;
    float c[kEsPolyCoefs];
};

// Smallest even 2,3,5-smooth n >= max(2M, 2w).
inline int fine_size(int M, int w) {
    int n = 2 * M > 2 * w ? 2 * M : 2 * w;
    for (;; ++n) {
        if (n % 2) continue;
        int m = n;
        for (int p : {2, 3, 5})
            while (m % p == 0) m /= p;
        if (m == 1) return n;
    }
}

// Gauss-Legendre nodes/weights on [-1, 1] (Newton on the Legendre recurrence).
inline void gauss_legendre(int nq, std::vector<double>& x, std::vector<double>& wq) {
    x.resize(nq);
    wq.resize(nq);
    for (int i = 0; i < nq; ++i) {
        double z = cos(M_PI * (i + 0.75) / (nq + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = 0.0;
            for (int j = 1; j <= nq; ++j) {
                const double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
            }
            const double dp = nq * (z * p0 - p1) / (z * z - 1.0);
            const double dz = p0 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = 0.0;
        for (int j = 1; j <= nq; ++j) {
            const double p2 = p1;
            p1 = p0;
            p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
        }
        const double dp = nq * (z * p0 - p1) / (z * z - 1.0);
        x[i] = z;
        wq[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

// phihat(xi) = int_{-w/2}^{w/2} phi(z) cos(xi z) dz.
inline double es_phihat(double xi, int w, double beta, const std::vector<double>& gx,
                        const std::vector<double>& gw) {
    double s = 0.0;
    for (size_t i = 0; i < gx.size(); ++i) {
        const double z = gx[i] * 0.5 * w;
        s += gw[i] * 0.5 * w * es_phi_host(z, w, beta) * cos(xi * z);
    }
    return s;
}

// Per-position deconvolution weights of an n-point line holding M modes: position
// p < ceil(M/2) holds k = p, position p >= n - floor(M/2) holds k = p - n; the
// weight is 1 / phihat(2 pi k / n).  Unused positions get 0.
inline std::vector<float> deconv_weights(int M, int n, int w, double beta) {
    std::vector<double> gx, gw;
    gauss_legendre(2 * w + 64, gx, gw);
    std::vector<float> out(n, 0.0f);
    const int kneg = M / 2, kpos = M - M / 2;
    for (int p = 0; p < n; ++p) {
        int k;
        if (p < kpos)
            k = p;
        else if (p >= n - kneg)
            k = p - n;
        else
            continue;
        out[p] = (float)(1.0 / es_phihat(2.0 * M_PI * k / n, w, beta, gx, gw));
    }
    return out;
}

}  // namespace cmtip
