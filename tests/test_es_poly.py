"""Host-side ES kernel tables (cmtip_amd/csrc/es.h) compiled with g++ and checked on
the CPU against the kernel's definition:

* es_poly: the per-tap polynomials every interpolation and spreading kernel evaluates
  instead of exp/sqrt.  Each tap of every width w <= 8 is rebuilt from the stored
  even / odd parts (mirror pairs E + yO and E - yO, the middle tap of odd w) and
  compared with phi(a - tau) on a dense grid of offsets; a swapped mirror sign, a
  shifted tap or a misplaced coefficient block gives errors of order 0.1.
* fine_size: the smallest even 2,3,5-smooth size >= max(2M, 2w), by brute force.
* es_phihat: the kernel's Fourier transform against scipy quadrature.

The ES kernel is exp(beta (sqrt(1 - (2z/w)^2) - 1)) on |z| <= w/2 with
beta = 2.30 w (Barnett, Magland & af Klinteberg 2019), the kernel finufft uses for the
reference's nufft3d1 / nufft3d2 (cmtip/nufft/forward.py:40, adjoint.py:33).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ES_H = os.path.join(os.path.dirname(__file__), "..", "cmtip_amd", "csrc", "es.h")

WRAPPER = r"""
#include "es.h"
extern "C" int es_poly_c(int w, double beta, float* c) { return cmtip::es_poly(w, beta, c) ? 1 : 0; }
extern "C" int fine_size_c(int M, int w) { return cmtip::fine_size(M, w); }
extern "C" double es_phihat_c(double xi, int w, double beta) {
    std::vector<double> gx, gw;
    cmtip::gauss_legendre(2 * w + 64, gx, gw);
    return cmtip::es_phihat(xi, w, beta, gx, gw);
}
extern "C" int es_poly_coefs() { return cmtip::kEsPolyCoefs; }
"""


@pytest.fixture(scope="module")
def es(tmp_path_factory):
    d = tmp_path_factory.mktemp("es")
    src = d / "es_wrap.cpp"
    src.write_text(WRAPPER)
    so = d / "libes.so"
    inc = os.path.abspath(os.path.dirname(ES_H))
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-I", inc, str(src), "-o", str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    lib.es_poly_c.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_float)]
    lib.fine_size_c.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.es_phihat_c.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double]
    lib.es_phihat_c.restype = ctypes.c_double
    return lib


def phi(z, w, beta):
    t = 2.0 * z / w
    s = 1.0 - t * t
    return np.where(s < 0, 0.0, np.exp(beta * (np.sqrt(np.maximum(s, 0.0)) - 1.0)))


def taps_from_poly(c, w, y):
    """The w tap weights at offsets y (tap a weighs phi(a - tau), tau = y + (w - 1)/2),
    rebuilt from the stored layout: per mirror pair a < w/2 the even coefficients
    (y^0, y^2, ...) then the odd ones (y^1, y^3, ...), then the middle tap's even part."""
    ne, no = (w + 1) // 2 + 1, (w + 2) // 2
    y2 = y * y
    out = np.zeros((w, y.size))
    k = 0
    for a in range(w // 2):
        E = sum(float(c[k + j]) * y2 ** j for j in range(ne))
        O = sum(float(c[k + ne + j]) * y2 ** j for j in range(no))
        out[a] = E + y * O
        out[w - 1 - a] = E - y * O
        k += ne + no
    if w & 1:
        out[w // 2] = sum(float(c[k + j]) * y2 ** j for j in range(ne))
    return out


@pytest.mark.parametrize("w", [2, 3, 4, 5, 6, 7, 8])
def test_es_poly_taps_match_phi(es, w):
    beta = 2.30 * w
    n = es.es_poly_coefs()
    c = (ctypes.c_float * n)()
    assert es.es_poly_c(w, beta, c) == 1
    y = np.linspace(-0.5, 0.5, 4001)[1:]  # (-1/2, 1/2]
    tau = y + 0.5 * (w - 1)
    exact = np.stack([phi(a - tau, w, beta) for a in range(w)])
    err = np.abs(taps_from_poly(c, w, y) - exact).max()
    eps = 10.0 ** (-(w - 1))  # the tolerance this width serves (w = ceil(log10(1/eps)) + 1)
    # measured: below eps / 7 for w <= 7 and eps / 3.7 at w = 8 (float32 coefficients)
    assert err <= eps / (3 if w == 8 else 5), (w, err)
    if w == 7:
        assert err <= 1e-7, err


def test_es_poly_refuses_wide_kernels(es):
    c = (ctypes.c_float * es.es_poly_coefs())()
    for w in (1, 9, 12, 16):
        assert es.es_poly_c(w, 2.3 * w, c) == 0
        assert all(v == 0.0 for v in c)


def test_es_poly_mirror_sign(es):
    """The odd part flips sign between a tap and its mirror: at y = +1/2 (tau = w/2)
    the first tap sits at the kernel's edge (phi(-w/2) = exp(-beta)), the last at
    phi(w/2 - 1); with the signs swapped they trade places."""
    w, beta = 7, 2.30 * 7
    c = (ctypes.c_float * es.es_poly_coefs())()
    es.es_poly_c(w, beta, c)
    t = taps_from_poly(c, w, np.array([0.5]))[:, 0]
    assert t[0] < 1e-6
    assert abs(t[-1] - phi(2.5, w, beta)) < 1e-7 and t[-1] > 1e-3
    # the same comparison as test_es_poly_taps_match_phi fails when the reconstruction
    # swaps the mirror signs (taps read as E - yO and E + yO)
    y = np.linspace(-0.5, 0.5, 4001)[1:]
    tau = y + 0.5 * (w - 1)
    exact = np.stack([phi(a - tau, w, beta) for a in range(w)])
    swapped = taps_from_poly(c, w, -y)
    assert np.abs(swapped - exact).max() > 1e-2


def fine_size_brute(M, w):
    n = max(2 * M, 2 * w)
    while True:
        if n % 2 == 0:
            m = n
            for p in (2, 3, 5):
                while m % p == 0:
                    m //= p
            if m == 1:
                return n
        n += 1


def test_fine_size(es):
    assert es.fine_size_c(81, 7) == 162
    assert es.fine_size_c(41, 7) == 90  # 82, 84 (7), 86 (43), 88 (11) are not 5-smooth
    assert es.fine_size_c(151, 7) == 320
    for M in range(1, 200):
        for w in (2, 7, 16):
            assert es.fine_size_c(M, w) == fine_size_brute(M, w), (M, w)


def test_es_phihat_vs_quadrature(es):
    from scipy.integrate import quad
    for w in (4, 7, 10):
        beta = 2.30 * w
        for xi in (0.0, 0.3, 1.0, 2.5):
            ref, _ = quad(lambda z: phi(z, w, beta) * np.cos(xi * z), -w / 2, w / 2, epsabs=1e-15, epsrel=1e-13,
                          limit=400, points=[-0.99 * w / 2, 0.99 * w / 2])
            got = es.es_phihat_c(xi, w, beta)
            # Gauss-Legendre on 2w + 64 nodes; phi's square-root edge limits it to
            # ~1e-5 eps relative (2.5e-9 at w = 4), far below what deconvolution needs
            eps = 10.0 ** (-(w - 1))
            assert abs(got - ref) <= max(1e-4 * eps, 1e-12) * abs(ref), (w, xi, got, ref)
