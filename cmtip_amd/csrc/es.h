// Exponential-of-semicircle (ES) spreading kernel (Barnett, Magland & af Klinteberg,
// SIAM J. Sci. Comput. 41(5) C479, 2019), upsampling factor 2.
//
// In fine-grid units z (|z| <= w/2):  phi(z) = exp(beta * (sqrt(1 - (2z/w)^2) - 1)).
// A point at fine-grid coordinate t = x * n / (2 pi) touches the w nodes
// l0, ..., l0 + w - 1 with l0 = ceil(t - w/2) (wrapped mod n).
#pragma once
#include <math.h>

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
