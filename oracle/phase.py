"""Phasing of an oversampled autocorrelation in FP64 (TEST INFRASTRUCTURE ONLY).

Restates cmtip/phasing/phase.py:11-219 with numpy complex128 FFTs and scipy's
gaussian_filter: Error Reduction (ER, :83-99), Hybrid Input-Output (HIO, :102-119),
the amplitude projection (step_phase, :122-144), shrink-wrap (:147-162), the initial
support (create_support_, :47-64), recentring (:11-44) and the phase driver
(:165-219).  Differences from the reference text: ``np.infty`` (:203, removed in
numpy 2) is ``np.inf``, and the random start of :201 takes a ``rng`` (default: numpy's
global generator, as the reference uses).
"""
import numpy as np
from scipy.ndimage import gaussian_filter


def center_of_mass(rho_, hkl_, M):
    """phase.py:11-23."""
    rho_ = np.abs(rho_)
    num = (rho_ * hkl_).sum(axis=(1, 2, 3))
    den = rho_.sum()
    return np.round(num / den * M / 2)


def hkl_grid(M):
    """The ifftshifted cell-centre coordinates in [-1, 1) of recenter (phase.py:34-38)."""
    ls = np.linspace(-1, 1, M + 1)
    ls = (ls[:-1] + ls[1:]) / 2
    hkl_list = np.meshgrid(ls, ls, ls, indexing="ij")
    return np.stack([np.fft.ifftshift(coord) for coord in hkl_list])


def recenter(rho_, support_, M):
    """phase.py:26-44, in place; returns the shift vector."""
    vect = center_of_mass(rho_, hkl_grid(M), M)
    for i in range(3):
        shift = int(vect[i])
        rho_[:] = np.roll(rho_, -shift, i)
        support_[:] = np.roll(support_, -shift, i)
    return vect


def create_support_(ac_, M, Mquat, generation=0):
    """phase.py:47-64."""
    sl = slice(Mquat, -Mquat)
    square_support = np.zeros((M, M, M), dtype=np.bool_)
    square_support[sl, sl, sl] = 1
    square_support_ = np.fft.ifftshift(square_support)
    thresh_support_ = ac_ > 1e-2 * ac_.max()
    return np.logical_and(square_support_, thresh_support_)


def step_phase(rho_, amplitudes_, amp_mask_, support_):
    """phase.py:122-144."""
    rho_hat_ = np.fft.fftn(rho_)
    phases_ = np.angle(rho_hat_)
    rho_hat_mod_ = np.where(amp_mask_, amplitudes_ * np.exp(1j * phases_), rho_hat_)
    rho_mod_ = np.fft.ifftn(rho_hat_mod_).real
    support_star_ = np.logical_and(support_, rho_mod_ > 0)
    return rho_mod_, support_star_


def ER(rho_, amplitudes_, amp_mask_, support_, rho_max):
    """phase.py:83-99, in place."""
    rho_mod_, support_star_ = step_phase(rho_, amplitudes_, amp_mask_, support_)
    rho_[:] = np.where(support_star_, rho_mod_, 0)
    i_overmax = rho_mod_ > rho_max
    rho_[i_overmax] = rho_max


def HIO(beta, rho_, amplitudes_, amp_mask_, support_, rho_max):
    """phase.py:102-119, in place."""
    rho_mod_, support_star_ = step_phase(rho_, amplitudes_, amp_mask_, support_)
    rho_[:] = np.where(support_star_, rho_mod_, rho_ - beta * rho_mod_)
    i_overmax = rho_mod_ > rho_max
    rho_[i_overmax] += 2 * beta * rho_mod_[i_overmax] - rho_max


def shrink_wrap(cutoff, sigma, rho_, support_):
    """phase.py:147-162, in place."""
    rho_abs_ = np.absolute(rho_)
    rho_gauss_ = gaussian_filter(rho_abs_, mode="wrap", sigma=sigma, truncate=2)
    support_[:] = rho_gauss_ > rho_abs_.max() * cutoff
    return rho_gauss_


def amplitudes_of(ac):
    """The filtered autocorrelation and its amplitudes (phase.py:186-195)."""
    M = ac.shape[0]
    ac_filt = gaussian_filter(np.maximum(ac.real, 0), mode="constant", sigma=1, truncate=2)
    ac_filt_ = np.fft.ifftshift(ac_filt)
    intensities_ = np.abs(np.fft.fftn(ac_filt_))
    amplitudes_ = np.sqrt(intensities_)
    amp_mask_ = np.ones((M, M, M), dtype=np.bool_)
    amp_mask_[0, 0, 0] = 0
    return ac_filt_, amplitudes_, amp_mask_


def phased_ac(rho_):
    """phase.py:214-217."""
    intensities_phased_ = np.abs(np.fft.fftn(rho_)) ** 2
    ac_phased_ = np.abs(np.fft.ifftn(intensities_phased_))
    return np.fft.fftshift(ac_phased_)


def phase(generation, ac, support_=None, rho_=None, n_iterations=10, nER=50, nHIO=25, rng=None):
    """phase.py:165-219: returns (ac_phased, support_, rho_)."""
    Mquat = int((ac.shape[0] - 1) / 4)
    M = ac.shape[0]
    ac_filt_, amplitudes_, amp_mask_ = amplitudes_of(np.asarray(ac, dtype=np.float64))
    if support_ is None:
        support_ = create_support_(ac_filt_, M, Mquat, generation)
    if rho_ is None:
        r = (rng.random(support_.shape) if rng is not None else np.random.rand(*support_.shape))
        rho_ = support_ * r
    rho_max = np.inf
    for i in range(n_iterations):
        for _ in range(nER):
            ER(rho_, amplitudes_, amp_mask_, support_, rho_max)
        for _ in range(nHIO):
            HIO(0.3, rho_, amplitudes_, amp_mask_, support_, rho_max)
        for _ in range(nER):
            ER(rho_, amplitudes_, amp_mask_, support_, rho_max)
        shrink_wrap(5e-2, 1, rho_, support_)
    for _ in range(nER):
        ER(rho_, amplitudes_, amp_mask_, support_, rho_max)
    recenter(rho_, support_, M)
    return phased_ac(rho_), support_, rho_
