"""Conversion between caller arrays (numpy or torch) and device tensors.

The reference functions take and return numpy arrays.  Here every function also
accepts torch tensors; results come back as numpy when the caller passed numpy, and
as device tensors when the caller passed tensors, so a GPU-resident pipeline never
copies through the host.
"""
import numpy as np
import torch

from ._lib import device


def is_torch(*xs):
    return any(isinstance(x, torch.Tensor) for x in xs)


def to_dev(x, dtype):
    """Device tensor of the given torch dtype (no copy when already there)."""
    if isinstance(x, torch.Tensor):
        return x.to(device=device(), dtype=dtype).contiguous()
    a = np.asarray(x)
    if dtype in (torch.complex64, torch.complex128):
        a = a.astype(np.complex64 if dtype == torch.complex64 else np.complex128, copy=False)
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device(), dtype=dtype)


def out(t, as_torch):
    return t if as_torch else t.cpu().numpy()


def max_abs(*xs):
    """max |x| over arrays or tensors, as a Python float (one host sync for tensors)."""
    m = 0.0
    for x in xs:
        if isinstance(x, torch.Tensor):
            if x.numel():
                m = max(m, float(x.detach().abs().max()))
        else:
            a = np.asarray(x)
            if a.size:
                m = max(m, float(np.max(np.abs(a))))
    return m


def all_real(x):
    """np.all(np.isreal(x)) for arrays and tensors."""
    if isinstance(x, torch.Tensor):
        if not torch.is_complex(x):
            return True
        return bool((x.imag == 0).all())
    return bool(np.all(np.isreal(x)))
