"""Phasing of an oversampled autocorrelation on the GPU (cmtip/phasing/phase.py).

Same functions, arguments and defaults as the reference; the work runs in the HIP
library (``cmtip_phaser_*`` / ``cmtip_phase_*``): every ER / HIO step is five FFT axis
passes with the amplitude projection and the density update folded in, the Gaussian
filters of shrink-wrap and of the amplitude set-up run in FP64, densities and FFTs in
FP32 / complex64.  Arrays whose names end in an underscore are in FFT position order
(DC at index 0), as in the reference (phase.py:4-9).  In-place functions update the
caller's array (a numpy array is copied to the device and back).  Supports are bool
numpy arrays or float 0 / 1 device tensors.
"""
import ctypes

import numpy as np
import torch

from .. import _arrays, _lib
from .._lib import check, ptr, stream

ER_MODE, HIO_MODE, PLAIN_MODE = 1, 2, 0


class Phaser:
    """A device phaser for M^3 volumes: amplitudes and an FFT scratch volume."""

    def __init__(self, M):
        h = ctypes.c_void_p()
        check(_lib.lib().cmtip_phaser_create(int(M), ctypes.byref(h)), "cmtip_phaser_create")
        self._h = h
        self.M = int(M)

    def setup(self, ac, Mquat=None):
        """Amplitudes from the centred autocorrelation (phase.py:186-195); returns
        (ac_filt_, support_) with support_ = create_support_ when Mquat is given."""
        M = self.M
        acd = _arrays.to_dev(np.real(ac) if not isinstance(ac, torch.Tensor) else ac.real, torch.float32)
        filt = torch.empty((M,) * 3, dtype=torch.float32, device=acd.device)
        sup = torch.empty((M,) * 3, dtype=torch.float32, device=acd.device) if Mquat is not None else None
        check(_lib.lib().cmtip_phase_setup(self._h, ptr(acd), ptr(filt), ptr(sup), int(Mquat or 0), stream()),
              "cmtip_phase_setup")
        return filt, sup

    def set_amplitudes(self, amplitudes_, amp_mask_=None):
        a = _arrays.to_dev(amplitudes_, torch.float32)
        if amp_mask_ is not None:
            a = torch.where(_arrays.to_dev(amp_mask_, torch.bool), a, torch.full_like(a, -1.0))
        check(_lib.lib().cmtip_phase_set_amplitudes(self._h, ptr(a.contiguous()), stream()),
              "cmtip_phase_set_amplitudes")

    def amplitudes(self):
        a = torch.empty((self.M,) * 3, dtype=torch.float32, device=_lib.device())
        check(_lib.lib().cmtip_phase_get_amplitudes(self._h, ptr(a), stream()), "cmtip_phase_get_amplitudes")
        return a

    def iterate(self, rho_, support_, mode, n, beta=0.0, rho_max=np.inf):
        check(_lib.lib().cmtip_phase_iterate(self._h, ptr(rho_), ptr(support_), int(mode), int(n), float(beta),
                                             float(rho_max), stream()), "cmtip_phase_iterate")

    def shrink_wrap(self, rho_, support_, cutoff, sigma):
        check(_lib.lib().cmtip_phase_shrink_wrap(self._h, ptr(rho_), float(cutoff), float(sigma), ptr(support_),
                                                 stream()), "cmtip_phase_shrink_wrap")

    def recenter(self, rho_, support_):
        sh = (ctypes.c_int * 3)()
        check(_lib.lib().cmtip_phase_recenter(self._h, ptr(rho_), ptr(support_), sh, stream()),
              "cmtip_phase_recenter")
        return np.array(sh[:], dtype=np.float64)

    def ac(self, rho_):
        out = torch.empty((self.M,) * 3, dtype=torch.float32, device=rho_.device)
        check(_lib.lib().cmtip_phase_ac(self._h, ptr(rho_), ptr(out), stream()), "cmtip_phase_ac")
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._LIB is not None:
            _lib._LIB.cmtip_phaser_destroy(h)
            self._h = None


_PHASERS = {}


def get_phaser(M):
    key = (int(M), torch.cuda.current_device())
    p = _PHASERS.get(key)
    if p is None:
        p = Phaser(M)
        _PHASERS[key] = p
    return p


def _vol(x):
    return _arrays.to_dev(x, torch.float32)


def _writeback(dst, src):
    """Copy a device result into the caller's array in place."""
    if isinstance(dst, torch.Tensor):
        if dst.data_ptr() != src.data_ptr():
            dst.copy_(src.to(dst.dtype))
    else:
        dst[...] = src.cpu().numpy().astype(dst.dtype)


def center_of_mass(rho_, hkl_, M):
    """Centre of mass of |rho_| in half-cells (phase.py:11-23).  Restated with torch on
    the caller's arrays (FP64); the recentring of ``phase`` uses the library's kernels."""
    r = torch.as_tensor(np.asarray(rho_) if not isinstance(rho_, torch.Tensor) else rho_).abs().double()
    h = torch.as_tensor(np.asarray(hkl_) if not isinstance(hkl_, torch.Tensor) else hkl_,
                        device=r.device).double()
    num = (r * h).sum(dim=(1, 2, 3))
    den = r.sum()
    return np.round((num / den * M / 2).cpu().numpy())


def recenter(rho_, support_, M):
    """Roll rho_ and support_ (in place) so the centre of mass sits at the origin
    (phase.py:26-44)."""
    ph = get_phaser(M)
    r, s = _vol(rho_), _vol(support_)
    ph.recenter(r, s)
    _writeback(rho_, r)
    _writeback(support_, s.bool() if not isinstance(support_, torch.Tensor) else s)


def create_support_(ac_, M, Mquat, generation):
    """Central-cube and threshold support of the filtered autocorrelation (phase.py:47-64);
    bool, FFT position order.  ac_ is in FFT position order, as in the reference."""
    a = _vol(ac_)
    cube = torch.zeros((M,) * 3, dtype=torch.bool, device=a.device)
    if Mquat > 0:
        cube[Mquat:-Mquat, Mquat:-Mquat, Mquat:-Mquat] = True
    sq = torch.roll(cube, shifts=(-(M // 2),) * 3, dims=(0, 1, 2))  # ifftshift
    s = sq & (a > 1e-2 * a.max())
    return s if isinstance(ac_, torch.Tensor) else s.cpu().numpy()


def step_phase(rho_, amplitudes_, amp_mask_, support_):
    """(rho_mod_, support_star_) of one amplitude projection (phase.py:122-144)."""
    M = rho_.shape[0]
    ph = get_phaser(M)
    ph.set_amplitudes(amplitudes_, amp_mask_)
    r = _vol(rho_).clone()
    ph.iterate(r, None, PLAIN_MODE, 1)
    star = (_vol(support_) > 0) & (r > 0)
    if isinstance(rho_, torch.Tensor):
        return r, star
    return r.cpu().numpy().astype(np.float64), star.cpu().numpy()


def ER(rho_, amplitudes_, amp_mask_, support_, rho_max):
    """One error-reduction step on rho_ in place (phase.py:83-99)."""
    ER_loop(1, rho_, amplitudes_, amp_mask_, support_, rho_max)


def HIO(beta, rho_, amplitudes_, amp_mask_, support_, rho_max):
    """One hybrid input-output step on rho_ in place (phase.py:102-119)."""
    HIO_loop(1, beta, rho_, amplitudes_, amp_mask_, support_, rho_max)


def ER_loop(n_loops, rho_, amplitudes_, amp_mask_, support_, rho_max):
    """n_loops ER steps (phase.py:67-72)."""
    ph = get_phaser(rho_.shape[0])
    ph.set_amplitudes(amplitudes_, amp_mask_)
    r = _vol(rho_)
    ph.iterate(r, _vol(support_), ER_MODE, n_loops, 0.0, rho_max)
    _writeback(rho_, r)


def HIO_loop(n_loops, beta, rho_, amplitudes_, amp_mask_, support_, rho_max):
    """n_loops HIO steps (phase.py:75-80)."""
    ph = get_phaser(rho_.shape[0])
    ph.set_amplitudes(amplitudes_, amp_mask_)
    r = _vol(rho_)
    ph.iterate(r, _vol(support_), HIO_MODE, n_loops, beta, rho_max)
    _writeback(rho_, r)


def shrink_wrap(cutoff, sigma, rho_, support_):
    """support_ = gaussian_filter(|rho_|, 'wrap', sigma, truncate 2) > max|rho_| * cutoff,
    in place (phase.py:147-162)."""
    ph = get_phaser(rho_.shape[0])
    s = torch.empty(tuple(rho_.shape), dtype=torch.float32, device=_lib.device())
    ph.shrink_wrap(_vol(rho_), s, cutoff, sigma)
    _writeback(support_, s.bool() if not isinstance(support_, torch.Tensor) else s)


def phase(generation, ac, support_=None, rho_=None, n_iterations=10, nER=50, nHIO=25):
    """Cycles of ER, HIO, ER and shrink-wrap from the autocorrelation, then nER ER steps
    and recentring (phase.py:165-219).  Returns (ac_phased, support_, rho_): centred
    float64 autocorrelation of the phased density, bool support, float64 density for
    numpy input (float32 device tensors, the support as float 0 / 1, for tensor input).
    Without rho_ the start is support_ * np.random.rand (numpy's global generator, as
    the reference draws it, :200-201)."""
    as_torch = _arrays.is_torch(ac, support_, rho_)
    M = ac.shape[0]
    Mquat = int((M - 1) / 4)
    ph = get_phaser(M)
    _, sup0 = ph.setup(ac, Mquat)
    sup = sup0 if support_ is None else _vol(support_).clone()
    if rho_ is None:
        start = sup.cpu().numpy() * np.random.rand(*sup.shape)
        rho = _vol(start)
    else:
        rho = _vol(rho_).clone()
    rho_max = np.inf
    for _ in range(n_iterations):
        ph.iterate(rho, sup, ER_MODE, nER, 0.0, rho_max)
        ph.iterate(rho, sup, HIO_MODE, nHIO, 0.3, rho_max)
        ph.iterate(rho, sup, ER_MODE, nER, 0.0, rho_max)
        ph.shrink_wrap(rho, sup, 5e-2, 1)
    ph.iterate(rho, sup, ER_MODE, nER, 0.0, rho_max)
    ph.recenter(rho, sup)
    acp = ph.ac(rho)
    if as_torch:
        return acp, sup, rho
    return (acp.cpu().numpy().astype(np.float64), sup.cpu().numpy() > 0, rho.cpu().numpy().astype(np.float64))
