"""FP64 oracle for the 3-D NUFFTs on the cmtip hot path (TEST INFRASTRUCTURE ONLY).

Conventions restated from the reference wrappers (which call finufft):
  forward / type 2 (cmtip/nufft/forward.py:13-42):
      c_j = sum_k f_k exp(-i k.x_j), k_d = i_d - floor(M/2)
      support multiplies the grid in place first (forward.py:34-35)
  adjoint / type 1 (cmtip/nufft/adjoint.py:13-43):
      f_k = sum_j c_j exp(+i k.x_j); then * support, then .real, then / M**3
  axis order: H <-> array axis 0, K <-> axis 1, L <-> axis 2 (forward.py:40,
  adjoint.py:33; the CPU path is normative).

Routes:
  nudft2 / nudft1        exact sums in C (oracle/nufft_oracle.c), OpenMP
  nudft2_np / nudft1_np  exact sums in numpy (separable einsum), small sizes only
  nufft2 / nufft1        ES-kernel fast route: C gridding + numpy FP64 FFT
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    """Load oracle/liboracle_nufft.so (built by oracle/Makefile)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_nufft.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-C", _HERE], check=True, capture_output=True)
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.c_int64
        ci = ctypes.c_int
        cd = ctypes.c_double
        L.oracle_interp.argtypes = [ci, ci, cd, dp, dp, i64, dp, dp, dp, dp, dp]
        L.oracle_spread.argtypes = [ci, ci, cd, i64, dp, dp, dp, dp, dp, dp, dp]
        L.oracle_nudft2.argtypes = [ci, dp, dp, i64, dp, dp, dp, dp, dp]
        L.oracle_nudft1.argtypes = [ci, i64, dp, dp, dp, dp, dp, dp, dp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_nudft1_modes.argtypes = [i64, dp, dp, dp, dp, dp, ci, ip, ci, ip, ci, ip, dp, dp]
        L.oracle_num_threads.restype = ci
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def modes(M):
    """Mode numbers k = i - floor(M/2) for array index i (finufft modeord 0)."""
    return np.arange(M) - M // 2


def es_params(eps):
    """ES kernel width w and shape beta at upsampling 2 (Barnett et al. 2019, sec. 3)."""
    w = int(np.ceil(np.log10(1.0 / eps))) + 1
    w = max(2, min(16, w))
    return w, 2.30 * w


def fine_size(M, w):
    """Smallest even 2,3,5-smooth n >= max(2M, 2w)."""
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


def es_phi(z, w, beta):
    t = 2.0 * np.asarray(z, dtype=np.float64) / w
    s = 1.0 - t * t
    return np.where(s >= 0, np.exp(beta * (np.sqrt(np.clip(s, 0, None)) - 1.0)), 0.0)


def es_phihat(xi, w, beta, nq=256):
    """phihat(xi) = int_{-w/2}^{w/2} phi(z) cos(xi z) dz by Gauss-Legendre quadrature."""
    x, wq = np.polynomial.legendre.leggauss(nq)
    z = x * w / 2.0
    wq = wq * w / 2.0
    return (wq * es_phi(z, w, beta) * np.cos(np.outer(np.asarray(xi, dtype=np.float64), z))).sum(1)


# ---------------------------------------------------------------- exact sums

def nudft2(f, x, y, z):
    """Exact type-2 sum (C, FP64).  f: (M,M,M) complex or real; returns complex128 (N,)."""
    f = np.asarray(f)
    M = f.shape[0]
    fr, fi = _f64(f.real), _f64(f.imag if np.iscomplexobj(f) else np.zeros_like(f.real))
    x, y, z = _f64(x).ravel(), _f64(y).ravel(), _f64(z).ravel()
    N = x.size
    cr, ci = np.empty(N), np.empty(N)
    lib().oracle_nudft2(M, _p(fr), _p(fi), N, _p(x), _p(y), _p(z), _p(cr), _p(ci))
    return cr + 1j * ci


def nudft1(c, x, y, z, M):
    """Exact type-1 sum (C, FP64).  Returns complex128 (M,M,M)."""
    c = np.asarray(c).ravel()
    cr = _f64(c.real)
    ci = _f64(c.imag if np.iscomplexobj(c) else np.zeros_like(cr))
    x, y, z = _f64(x).ravel(), _f64(y).ravel(), _f64(z).ravel()
    fr, fi = np.empty((M, M, M)), np.empty((M, M, M))
    lib().oracle_nudft1(M, x.size, _p(x), _p(y), _p(z), _p(cr), _p(ci), _p(fr), _p(fi))
    return fr + 1j * fi


def nudft1_modes(c, x, y, z, ka, kb, kc):
    """Exact type-1 sums (C, FP64) at the tensor product of mode numbers ka x kb x kc
    (k = array index - floor(M/2)); returns complex128 (len(ka), len(kb), len(kc)).
    Linear in the points, so usable at tens of millions of points."""
    c = np.asarray(c).ravel()
    cr = _f64(c.real)
    ci = _f64(c.imag) if np.iscomplexobj(c) else None
    x, y, z = _f64(x).ravel(), _f64(y).ravel(), _f64(z).ravel()
    ks = [np.ascontiguousarray(k, dtype=np.int32) for k in (ka, kb, kc)]
    assert all(1 <= k.size <= 256 for k in ks)
    fr, fi = np.empty((ks[0].size, ks[1].size, ks[2].size)), np.empty((ks[0].size, ks[1].size, ks[2].size))
    ip = ctypes.POINTER(ctypes.c_int)
    lib().oracle_nudft1_modes(x.size, _p(x), _p(y), _p(z), _p(cr), _p(ci) if ci is not None else None,
                              ks[0].size, ks[0].ctypes.data_as(ip), ks[1].size, ks[1].ctypes.data_as(ip),
                              ks[2].size, ks[2].ctypes.data_as(ip), _p(fr), _p(fi))
    return fr + 1j * fi


def nudft2_np(f, x, y, z):
    """Exact type-2 sum in numpy (independent of the C code); small sizes only."""
    k = modes(f.shape[0])
    E = [np.exp(-1j * np.outer(np.asarray(v, np.float64).ravel(), k)) for v in (x, y, z)]
    return np.einsum("abc,ja,jb,jc->j", f, E[0], E[1], E[2], optimize=True)


def nudft1_np(c, x, y, z, M):
    """Exact type-1 sum in numpy (independent of the C code); small sizes only."""
    k = modes(M)
    E = [np.exp(1j * np.outer(np.asarray(v, np.float64).ravel(), k)) for v in (x, y, z)]
    return np.einsum("j,ja,jb,jc->abc", np.asarray(c).ravel(), E[0], E[1], E[2], optimize=True)


# ---------------------------------------------------------------- ES fast route

def nufft2(f, x, y, z, eps=1e-12):
    """Type 2 by the ES route in FP64: deconvolve+pad, FFT (sign -1), interpolate."""
    f = np.asarray(f)
    M = f.shape[0]
    w, beta = es_params(eps)
    n = fine_size(M, w)
    h = 2 * np.pi / n
    k = modes(M)
    p = 1.0 / es_phihat(k * h, w, beta)
    b = np.zeros((n, n, n), np.complex128)
    idx = k % n
    b[np.ix_(idx, idx, idx)] = f * p[:, None, None] * p[None, :, None] * p[None, None, :]
    g = np.fft.fftn(b)
    x, y, z = _f64(x).ravel(), _f64(y).ravel(), _f64(z).ravel()
    gr, gi = _f64(g.real), _f64(g.imag)
    N = x.size
    cr, ci = np.empty(N), np.empty(N)
    lib().oracle_interp(n, w, beta, _p(gr), _p(gi), N, _p(x), _p(y), _p(z), _p(cr), _p(ci))
    return cr + 1j * ci


def nufft1(c, x, y, z, M, eps=1e-12):
    """Type 1 by the ES route in FP64: spread, FFT (sign +1), crop+deconvolve."""
    c = np.asarray(c).ravel()
    w, beta = es_params(eps)
    n = fine_size(M, w)
    h = 2 * np.pi / n
    x, y, z = _f64(x).ravel(), _f64(y).ravel(), _f64(z).ravel()
    cr = _f64(c.real)
    ci = _f64(c.imag if np.iscomplexobj(c) else np.zeros_like(cr))
    br, bi = np.zeros((n, n, n)), np.zeros((n, n, n))
    lib().oracle_spread(n, w, beta, x.size, _p(x), _p(y), _p(z), _p(cr), _p(ci), _p(br), _p(bi))
    F = np.fft.ifftn(br + 1j * bi) * float(n) ** 3
    k = modes(M)
    p = 1.0 / es_phihat(k * h, w, beta)
    idx = k % n
    return F[np.ix_(idx, idx, idx)] * p[:, None, None] * p[None, :, None] * p[None, None, :]


# ---------------------------------------------------------------- reference semantics

def forward(ugrid, H_, K_, L_, support=None, use_recip_sym=True, eps=1e-12):
    """cmtip.nufft.forward_cpu semantics (forward.py:13-42) in FP64."""
    assert np.max(np.abs(np.array([H_, K_, L_]))) < 3 * np.pi
    if use_recip_sym:
        assert np.all(np.isreal(ugrid))
    if support is not None:
        ugrid *= support
    return nufft2(ugrid, H_, K_, L_, eps=eps)


def adjoint(nuvect, H_, K_, L_, M, use_recip_sym=True, support=None, eps=1e-12):
    """cmtip.nufft.adjoint_cpu semantics (adjoint.py:13-43) in FP64."""
    assert np.max(np.abs(np.array([H_, K_, L_]))) < 3 * np.pi
    ugrid = nufft1(nuvect, H_, K_, L_, M, eps=eps)
    if support is not None:
        ugrid = ugrid * support
    if use_recip_sym:
        ugrid = ugrid.real
    return ugrid / (M ** 3)


def num_threads():
    return lib().oracle_num_threads()
