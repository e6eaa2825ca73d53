"""Pins the FP64 NUFFT oracle: two independent exact sums, the reference's uniform-grid
identity (cmtip/tests/test_nufft.py:31-49), the adjoint identity, and the ES route
against the exact sums."""
import numpy as np
import pytest

from oracle import nufft as O
from tests.conftest import rel_l2


@pytest.mark.parametrize("M", [5, 6, 8])
def test_exact_sums_agree(M, rng):
    N = 300
    f = rng.standard_normal((M, M, M)) + 1j * rng.standard_normal((M, M, M))
    x, y, z = rng.uniform(-3 * np.pi + 1e-3, 3 * np.pi - 1e-3, (3, N))
    assert rel_l2(O.nudft2(f, x, y, z), O.nudft2_np(f, x, y, z)) < 1e-12
    c = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    assert rel_l2(O.nudft1(c, x, y, z, M), O.nudft1_np(c, x, y, z, M)) < 1e-12


@pytest.mark.parametrize("M", [7, 9])
def test_uniform_grid_identity(M, rng):
    # points at x = pi * (2(i - M//2)/M) (odd M): type 2 = fftshift(fftn(ifftshift(f)))
    ls = np.linspace(-1, 1, M + 1)
    ls = (ls[:-1] + ls[1:]) / 2
    H, K, L = (np.pi * g.ravel() for g in np.meshgrid(ls, ls, ls, indexing="ij"))
    f = rng.standard_normal((M, M, M))
    want = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(f)))
    assert rel_l2(O.nudft2(f, H, K, L), want.ravel()) < 1e-12
    assert rel_l2(O.nufft2(f, H, K, L, eps=1e-12), want.ravel()) < 1e-10
    # type 1 of the transform returns M^3 f (the adjoint identity on the uniform grid)
    back = O.nufft1(want.ravel(), H, K, L, M, eps=1e-12) / M ** 3
    assert rel_l2(back, f) < 1e-10


def test_adjoint_identity(rng):
    M, N = 9, 2000
    f = rng.standard_normal((M, M, M)) + 1j * rng.standard_normal((M, M, M))
    c = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    x, y, z = rng.uniform(-np.pi, np.pi, (3, N))
    lhs = np.vdot(O.nufft2(f, x, y, z), c)  # <A f, c>
    # A^H c with A the sign -1 transform is the sign +1 sum
    rhs = np.vdot(f.ravel(), O.nufft1(c, x, y, z, M).ravel())
    assert abs(lhs - rhs) / abs(lhs) < 1e-10


@pytest.mark.parametrize("M,eps,tol", [(8, 1e-12, 1e-10), (11, 1e-12, 1e-10), (11, 1e-6, 2e-5)])
def test_es_route_vs_exact(M, eps, tol, rng):
    N = 2000
    f = rng.standard_normal((M, M, M))
    x, y, z = rng.uniform(-3 * np.pi + 1e-3, 3 * np.pi - 1e-3, (3, N))
    assert rel_l2(O.nufft2(f, x, y, z, eps=eps), O.nudft2(f, x, y, z)) < tol
    c = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    assert rel_l2(O.nufft1(c, x, y, z, M, eps=eps), O.nudft1(c, x, y, z, M)) < tol


def test_planted_mistake_is_caught(rng):
    # a sign flip in the exponent (type 1 vs type 2 convention) must fail the check above
    M, N = 6, 200
    f = rng.standard_normal((M, M, M)) + 1j * rng.standard_normal((M, M, M))
    x, y, z = rng.uniform(-np.pi, np.pi, (3, N))
    wrong = O.nudft2_np(f, -x, -y, -z)
    assert rel_l2(wrong, O.nudft2(f, x, y, z)) > 1e-2


def test_es_params():
    assert O.es_params(1e-6) == (7, 2.30 * 7)
    assert O.fine_size(81, 7) == 162
    assert O.fine_size(41, 7) == 90
    assert O.fine_size(162, 7) == 324
