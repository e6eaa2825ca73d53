"""Pins the autocorrelation oracle: the Toeplitz identity against a brute-force
A^H A, the Fourier-corner projection against explicit DFT matrices, and the scipy CG
solve against a dense solve of the same operator."""
import numpy as np
import pytest

from oracle import ac, nufft
from tests.conftest import rel_l2


def _dft_matrix(x, M, sign):
    k = np.arange(M) - M // 2
    return np.exp(sign * 1j * np.outer(x, k))


@pytest.mark.parametrize("M", [5, 6])
def test_toeplitz_identity(M, rng):
    N = 400
    x, y, z = rng.uniform(-np.pi, np.pi, (3, N))
    # dense A (N x M^3) with the type-2 sign
    Ax, Ay, Az = (_dft_matrix(v, M, -1) for v in (x, y, z))
    A = np.einsum("ja,jb,jc->jabc", Ax, Ay, Az).reshape(N, -1)
    S = (rng.random((M, M, M)) > 0.2).astype(float)
    u = rng.standard_normal(M ** 3)
    want = (np.real(A.conj().T @ (A @ (S.ravel() * u))) / (2 * M) ** 3) * S.ravel()
    F = ac.kernel_spectrum(x, y, z, M, exact=True)
    got = ac.core_problem_convolution(u, M, F, 2 * M, S)
    assert rel_l2(got, want) < 1e-12
    # a larger even embedding gives the same operator
    F2 = ac.kernel_spectrum(x, y, z, M, L=2 * M + 4, exact=True)
    assert rel_l2(ac.core_problem_convolution(u, M, F2, 2 * M + 4, S), want) < 1e-12


@pytest.mark.parametrize("M", [5, 7, 8])
def test_fourier_reg_explicit(M, rng):
    anti = ac.gen_F_antisupport(M)
    S = (rng.random((M, M, M)) > 0.3).astype(float)
    u = rng.standard_normal(M ** 3)
    c = np.arange(M) - M // 2
    E = np.exp(-2j * np.pi * np.outer(c, c) / M)  # E[k, x] over centred k, x
    g = (S.ravel() * u).reshape(M, M, M)
    G = np.einsum("ka,lb,mc,abc->klm", E, E, E, g)
    R = np.einsum("ak,bl,cm,klm->abc", E.conj(), E.conj(), E.conj(), G * anti) / M ** 3
    want = (R.real * S).ravel()
    assert rel_l2(ac.fourier_reg(u, S, anti, M), want) < 1e-12


def test_antisupport_shape():
    for M in (5, 41, 81):
        m = ac.gen_F_antisupport(M)
        assert m[0, 0, 0] and not m[M // 2, M // 2, M // 2]
        assert np.all(m == m[::-1, ::-1, ::-1])


def test_cg_matches_dense_solve(rng):
    M = 5
    N = 600
    q = rng.standard_normal((N // 25, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    n_side = 5
    from cmtip_amd import synth
    s, ext = synth.detector(n_side)
    acm = synth.gaussian_ac(M, 5, seed=3)
    from oracle import geom
    x, y, z = geom.rotated_coords(q, s, ext)
    data = nufft.nudft2(acm.astype(float), x, y, z).real
    res, its, sup, d, F = ac.solve(q, s, ext, data.reshape(len(q), -1), M, maxiter=500, exact=True, rtol=1e-12)
    anti = ac.gen_F_antisupport(M)
    rl = M ** 3 / data.size / 1000
    W = np.stack([ac.core_problem_convolution(e, M, F, 2 * M, sup) + rl * e + 1e3 * ac.fourier_reg(e, sup, anti, M)
                  for e in np.eye(M ** 3)], axis=1)
    np.testing.assert_allclose(W, W.T, atol=1e-10 * np.abs(W).max())  # symmetric
    assert np.linalg.eigvalsh(W).min() > 0  # positive definite
    want = np.linalg.solve(W, d)
    assert rel_l2(res.ravel(), want) < 1e-8
