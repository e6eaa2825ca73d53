"""NUFFT plan objects over the C ABI (the cufinufft v1 plan the reference builds per
call, cmtip/nufft/forward.py:77-81 and cmtip/nufft/adjoint.py:73-76).

Plans own their fine grid and FFT scratch on the device.  ``get_plan`` keeps one plan
per (type, modes, sign, eps, device) so repeated calls reuse the allocations; plans
are not thread-safe.
"""
import ctypes

import torch

from .. import _lib
from .._lib import check, ptr, stream

DEFAULT_EPS = 1e-6


class NufftPlan:
    def __init__(self, nufft_type, n_modes, iflag, eps=DEFAULT_EPS):
        L = _lib.lib()
        h = ctypes.c_void_p()
        check(L.cmtip_nufft_makeplan(int(nufft_type), int(n_modes), int(iflag), float(eps), ctypes.byref(h)),
              "cmtip_nufft_makeplan")
        self._h = h
        self.type = int(nufft_type)
        self.n_modes = int(n_modes)
        self.iflag = int(iflag)
        self.eps = float(eps)
        self._pts = None
        w, n, beta = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(L.cmtip_nufft_info(h, ctypes.byref(w), ctypes.byref(n), ctypes.byref(beta)), "cmtip_nufft_info")
        self.w, self.n_fine, self.beta = w.value, n.value, beta.value

    def setpts(self, x, y, z):
        """x, y, z: float32 device tensors of equal length (kept alive by the plan)."""
        n = x.numel()
        assert y.numel() == n and z.numel() == n
        for t in (x, y, z):
            assert t.dtype == torch.float32 and t.is_cuda and t.is_contiguous()
        self._pts = (x, y, z)
        check(_lib.lib().cmtip_nufft_setpts(self._h, n, ptr(x), ptr(y), ptr(z), stream()), "cmtip_nufft_setpts")

    def execute(self, c, f, flags, support=None, scale=1.0):
        check(_lib.lib().cmtip_nufft_execute(self._h, ptr(c), ptr(f), int(flags), ptr(support), float(scale),
                                             stream()), "cmtip_nufft_execute")

    def slice_prepare(self, ac, support=None):
        check(_lib.lib().cmtip_slice_prepare(self._h, ptr(ac), ptr(support), stream()), "cmtip_slice_prepare")

    def slice_real(self, quats, s, P, recip_extent, out, out_bf16=None, moments=None, order=None):
        n_orient = quats.shape[0]
        check(_lib.lib().cmtip_slice_real(self._h, ptr(quats), n_orient, ptr(s), ptr(order), int(P),
                                          float(recip_extent), ptr(out), ptr(out_bf16), ptr(moments), stream()),
              "cmtip_slice_real")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._LIB is not None:
            _lib._LIB.cmtip_nufft_destroy(h)
            self._h = None


_CACHE = {}


def get_plan(nufft_type, n_modes, iflag, eps=DEFAULT_EPS):
    key = (int(nufft_type), int(n_modes), int(iflag), float(eps), torch.cuda.current_device())
    p = _CACHE.get(key)
    if p is None:
        p = NufftPlan(nufft_type, n_modes, iflag, eps)
        _CACHE[key] = p
    return p


def clear_plans():
    _CACHE.clear()
