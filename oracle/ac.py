"""The autocorrelation normal equations and their CG solve in FP64 (TEST
INFRASTRUCTURE ONLY).

Restates cmtip/autocorrelation/autocorrelation.py:32-248 with numpy complex128 FFTs
and scipy's cg (scipy/sparse/linalg/_isolve/iterative.py:305-422), using the FP64
NUFFT oracle for the two type-1 transforms.
"""
import numpy as np
from scipy.ndimage import gaussian_filter
from scipy.sparse.linalg import LinearOperator, cg

from . import geom, nufft


def core_problem_convolution(uvect, M, F_ugrid_conv_, M_ups, ac_support):
    """autocorrelation.py:32-63."""
    u = np.asarray(uvect).reshape((M,) * 3) * ac_support
    big = np.zeros((M_ups,) * 3, dtype=np.complex128)
    big[:M, :M, :M] = u
    conv = np.fft.fftshift(np.fft.ifftn(np.fft.fftn(np.fft.ifftshift(big)) * F_ugrid_conv_))
    return (conv[:M, :M, :M] * ac_support).real.flatten()


def gen_F_antisupport(M):
    """autocorrelation.py:66-85."""
    lu = np.linspace(-np.pi, np.pi, M)
    A, B, C = np.meshgrid(lu, lu, lu, indexing="ij")
    return np.around(np.sqrt(A ** 2 + B ** 2 + C ** 2), 4) > np.pi


def fourier_reg(uvect, support, F_antisupport, M):
    """autocorrelation.py:88-110."""
    g = np.asarray(uvect).reshape((M,) * 3) * support
    r = np.fft.fftshift(np.fft.ifftn(np.fft.fftn(np.fft.ifftshift(g)) * np.fft.ifftshift(F_antisupport)))
    return (r * support).real.flatten()


def kernel_spectrum(x, y, z, M, L=None, eps=1e-12, exact=False):
    """F_ugrid_conv_ (autocorrelation.py:141-148): spectrum of the real type 1 of ones at
    L (default 2M) modes, normalised by (2M)^3."""
    L = 2 * M if L is None else L
    ones = np.ones(len(x))
    K = (nufft.nudft1(ones, x, y, z, L) if exact else nufft.nufft1(ones, x, y, z, L, eps=eps)).real
    K = K / float(2 * M) ** 3
    return np.fft.fftn(np.fft.ifftshift(K))


def rhs(x, y, z, data, M, support, x0, alambda, rlambda, eps=1e-12, exact=False):
    """d = alambda * Re(support * T1(data)) / M^3 + rlambda * x0 (autocorrelation.py:162-171)."""
    c = np.asarray(data, dtype=np.float64).ravel()
    f = nufft.nudft1(c, x, y, z, M) if exact else nufft.nufft1(c, x, y, z, M, eps=eps)
    adb = (f * support).real / M ** 3
    return alambda * adb.flatten() + rlambda * np.asarray(x0, dtype=np.float64).ravel()


def support_of(ac_estimate):
    """autocorrelation.py:218-219."""
    return (gaussian_filter(ac_estimate, 0.5) > 1e-12).astype(np.float32)


def solve(quats, pixel_position_reciprocal, reciprocal_extent, slices_, M, ac_estimate=None, maxiter=100,
          rlambda=None, flambda=1e3, alambda=1.0, L=None, eps=1e-12, exact=False, rtol=1e-5):
    """solve_ac (autocorrelation.py:176-248) in complex128 with scipy cg.

    Returns (ac, iterations, support, d, F_ugrid_conv_)."""
    x, y, z = geom.rotated_coords(quats, pixel_position_reciprocal, reciprocal_extent)
    Mtot = M ** 3
    N = int(np.prod(np.shape(slices_)))
    if ac_estimate is None:
        support = np.ones((M,) * 3)
        x0 = np.zeros(Mtot)
    else:
        est = np.array(ac_estimate, dtype=np.float64)
        support = support_of(np.asarray(ac_estimate)).astype(np.float64)
        x0 = (est * support).ravel()
    if rlambda is None:
        rlambda = Mtot / N / 1000
    F = kernel_spectrum(x, y, z, M, L, eps, exact)
    Lm = F.shape[0]
    anti = gen_F_antisupport(M)

    def matvec(u):
        u = np.real(u)
        return (alambda * core_problem_convolution(u, M, F, Lm, support) + rlambda * u
                + flambda * fourier_reg(u, support, anti, M))

    W = LinearOperator(dtype=np.complex128, shape=(Mtot, Mtot), matvec=matvec)
    d = rhs(x, y, z, slices_, M, support, x0, alambda, rlambda, eps, exact)
    count = [0]

    def cb(_):
        count[0] += 1

    ret, info = cg(W, d.astype(np.complex128), x0=x0.astype(np.complex128), maxiter=maxiter, rtol=rtol,
                   callback=cb)
    return ret.real.reshape((M,) * 3), count[0], support, d, F
