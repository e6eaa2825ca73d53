"""The phasing oracle (oracle/phase.py) pinned by fixed points, brute-force sums and
constructed cases (no GPU)."""
import numpy as np
import pytest

from oracle import phase as P


def _density(M, seed=0):
    rng = np.random.default_rng(seed)
    rho = np.zeros((M,) * 3)
    c = M // 2
    r = max(1, M // 5)
    rho[c - r:c + r, c - r:c + r, c - r:c + r] = rng.random((2 * r,) * 3)
    return np.fft.ifftshift(rho)


def test_true_density_is_a_fixed_point_of_er_and_hio():
    # with the exact amplitudes of a non-negative density inside its support, the
    # projection returns the density itself, so ER and HIO leave it unchanged
    M = 12
    rho = _density(M)
    amp = np.abs(np.fft.fftn(rho))
    mask = np.ones_like(rho, bool)
    mask[0, 0, 0] = False
    sup = rho > 0
    for step in ("ER", "HIO"):
        r = rho.copy()
        if step == "ER":
            P.ER(r, amp, mask, sup, np.inf)
        else:
            P.HIO(0.3, r, amp, mask, sup, np.inf)
        np.testing.assert_allclose(r, rho, atol=1e-12)


def test_er_projects_onto_support_and_positivity():
    M = 10
    rng = np.random.default_rng(1)
    rho = rng.standard_normal((M,) * 3)
    amp = np.abs(np.fft.fftn(rng.random((M,) * 3)))
    mask = np.ones_like(rho, bool)
    mask[0, 0, 0] = False
    sup = rng.random((M,) * 3) > 0.5
    r = rho.copy()
    P.ER(r, amp, mask, sup, np.inf)
    assert np.all(r[~sup] == 0) and np.all(r >= 0)
    # the amplitudes of the projected density before the real-space step are the target
    mod, star = P.step_phase(rho, amp, mask, sup)
    np.testing.assert_array_equal(r, np.where(star, mod, 0))
    # a planted mistake (the DC amplitude replaced too) changes the result
    r2 = rho.copy()
    P.ER(r2, amp, np.ones_like(mask), sup, np.inf)
    assert not np.allclose(r2, r)


def test_hio_update_rule():
    M = 8
    rng = np.random.default_rng(2)
    rho = rng.standard_normal((M,) * 3)
    amp = np.abs(np.fft.fftn(rng.random((M,) * 3)))
    mask = np.ones_like(rho, bool)
    mask[0, 0, 0] = False
    sup = rng.random((M,) * 3) > 0.3
    mod, star = P.step_phase(rho, amp, mask, sup)
    r = rho.copy()
    P.HIO(0.3, r, amp, mask, sup, np.inf)
    want = np.where(star, mod, rho - 0.3 * mod)
    np.testing.assert_allclose(r, want, rtol=0, atol=0)
    # rho_max: ER clips, HIO adds 2 beta r - rho_max where r exceeds it
    r = rho.copy()
    P.ER(r, amp, mask, sup, 0.1)
    assert np.all(r[mod > 0.1] == 0.1)


def test_phased_ac_is_the_circular_autocorrelation():
    M = 6
    rho = np.random.default_rng(3).random((M,) * 3)
    got = P.phased_ac(rho)
    # brute force: a(d) = sum_x rho(x) rho(x + d) (circular), fftshifted
    want = np.zeros((M,) * 3)
    for d in np.ndindex(M, M, M):
        want[d] = np.sum(rho * np.roll(rho, tuple(-k for k in d), axis=(0, 1, 2)))
    np.testing.assert_allclose(got, np.fft.fftshift(want), rtol=1e-12)


@pytest.mark.parametrize("mode", ["wrap", "constant"])
def test_gaussian_filters_against_explicit_sums(mode):
    # the filters of shrink-wrap (wrap) and of the amplitudes (constant), sigma 1,
    # truncate 2: 5-tap weights exp(-x^2 / 2) normalised, one axis at a time
    from scipy.ndimage import gaussian_filter
    M = 7
    a = np.random.default_rng(4).random((M,) * 3)
    w = np.exp(-0.5 * np.arange(-2, 3) ** 2)
    w /= w.sum()
    want = a.copy()
    for ax in range(3):
        out = np.zeros_like(want)
        for k, wk in zip(range(-2, 3), w):
            if mode == "wrap":
                out += wk * np.roll(want, -k, axis=ax)
            else:
                sh = np.zeros_like(want)
                src = [slice(None)] * 3
                dst = [slice(None)] * 3
                if k >= 0:
                    src[ax], dst[ax] = slice(k, M), slice(0, M - k)
                else:
                    src[ax], dst[ax] = slice(0, M + k), slice(-k, M)
                sh[tuple(dst)] = want[tuple(src)]
                out += wk * sh
        want = out
    np.testing.assert_allclose(gaussian_filter(a, mode=mode, sigma=1, truncate=2), want, rtol=1e-12)
    if mode == "wrap":
        rho = a.copy()
        sup = np.zeros_like(a, bool)
        g = P.shrink_wrap(0.3, 1, rho, sup)
        np.testing.assert_allclose(g, want, rtol=1e-12)
        np.testing.assert_array_equal(sup, want > a.max() * 0.3)


def test_recenter_moves_a_blob_to_the_origin():
    M = 16
    rho = np.zeros((M,) * 3)
    rho[5, 9, 3] = 1.0  # FFT position order: cell (5, 9, 3)
    sup = rho > 0
    vect = P.recenter(rho, sup, M)
    # hkl_ of cell i along an axis is ls[(i + M/2) % M], ls = cell centres of [-1, 1)
    ls = (np.linspace(-1, 1, M + 1)[:-1] + np.linspace(-1, 1, M + 1)[1:]) / 2
    want = np.round(np.array([ls[(i + M // 2) % M] for i in (5, 9, 3)]) * M / 2)
    np.testing.assert_array_equal(vect, want)
    pos = np.argwhere(rho > 0)[0]
    np.testing.assert_array_equal(pos, [(i - int(v)) % M for i, v in zip((5, 9, 3), want)])
    assert sup[tuple(pos)]


def test_create_support_cube_and_threshold():
    M, Mquat = 13, 3
    ac = np.zeros((M,) * 3)
    ac[M // 2, M // 2, M // 2] = 1.0
    ac[M // 2 + 1, M // 2, M // 2] = 0.02
    ac[M // 2 + 1, M // 2 + 1, M // 2] = 0.005   # below 1e-2 of the maximum
    ac[0, 0, 0] = 0.5                             # outside the central cube
    s = P.create_support_(np.fft.ifftshift(ac), M, Mquat)
    want = np.zeros((M,) * 3, bool)
    want[M // 2, M // 2, M // 2] = want[M // 2 + 1, M // 2, M // 2] = True
    np.testing.assert_array_equal(s, np.fft.ifftshift(want))
    # Mquat = 0: slice(0, -0) is empty, so is the support
    assert not P.create_support_(np.fft.ifftshift(ac), M, 0).any()
