"""Forward (type 2) NUFFT: grid -> nonuniform points, isign -1.

Mirrors cmtip/nufft/forward.py:13-82.  Both names run the HIP path; ``forward_cpu``
exists so callers written against the reference keep working.
"""
import numpy as np
import torch

from .. import _arrays
from .._lib import CMTIP_GRID_REAL, CMTIP_PTS_REAL
from .plan import DEFAULT_EPS, get_plan


def _forward(ugrid, H_, K_, L_, support, use_recip_sym, eps, real_out=False):
    x = _arrays.to_dev(H_, torch.float32).reshape(-1)
    y = _arrays.to_dev(K_, torch.float32).reshape(-1)
    z = _arrays.to_dev(L_, torch.float32).reshape(-1)
    plan = get_plan(2, ugrid.shape[0], -1, eps)
    plan.setpts(x, y, z)
    # points must lie in finufft's permissible domain (forward.py:27, 64); the maximum is
    # recorded while setpts bins the points
    assert plan.points_max_abs() < 3 * np.pi
    # the reference applies the support in place on the caller's grid (forward.py:34-35, 66-67)
    if support is not None:
        if isinstance(ugrid, torch.Tensor):
            ugrid *= torch.as_tensor(support, device=ugrid.device, dtype=ugrid.dtype if not torch.is_complex(ugrid)
                                     else ugrid.real.dtype)
        else:
            ugrid *= support
    as_torch = _arrays.is_torch(ugrid, H_)
    shape = tuple(H_.shape)
    real_grid = use_recip_sym or not (torch.is_complex(ugrid) if isinstance(ugrid, torch.Tensor)
                                      else np.iscomplexobj(ugrid))
    if real_grid:
        src = ugrid.real if (isinstance(ugrid, torch.Tensor) and torch.is_complex(ugrid)) else (
            np.real(ugrid) if not isinstance(ugrid, torch.Tensor) else ugrid)
        f = _arrays.to_dev(src, torch.float32)
    else:
        f = _arrays.to_dev(ugrid, torch.complex64)
    M = f.shape[0]
    assert f.dim() == 3 and f.shape[1] == M and f.shape[2] == M, "ugrid must be (M, M, M)"
    flags = CMTIP_GRID_REAL if real_grid else 0
    if real_out:
        c = torch.empty(x.numel(), dtype=torch.float32, device=f.device)
        flags |= CMTIP_PTS_REAL
    else:
        c = torch.empty(x.numel(), dtype=torch.complex64, device=f.device)
    plan.execute(c, f, flags)
    return _arrays.out(c.reshape(shape), as_torch)


def forward_cpu(ugrid, H_, K_, L_, support, use_recip_sym, *, eps=DEFAULT_EPS):
    """Type-2 NUFFT of ugrid at (H_, K_, L_): c_j = sum_k f_k exp(-i k.x_j).

    Signature of cmtip/nufft/forward.py:13; returns complex64 of H_'s shape.  Runs on
    the GPU (the reference's finufft call at forward.py:40 is replaced, not kept).
    ``eps`` sets the ES kernel width (the reference's 1e-12 exceeds FP32 precision).
    """
    if use_recip_sym:
        assert _arrays.all_real(ugrid)
    return _forward(ugrid, H_, K_, L_, support, use_recip_sym, eps)


def forward_gpu(ugrid, H_, K_, L_, support=None, use_recip_sym=True, *, eps=DEFAULT_EPS):
    """Type-2 NUFFT; signature of cmtip/nufft/forward.py:45.  With use_recip_sym the
    real part of ugrid is transformed (forward.py:69-70)."""
    return _forward(ugrid, H_, K_, L_, support, use_recip_sym, eps)


def forward_real(ugrid, H_, K_, L_, support=None, *, eps=DEFAULT_EPS):
    """Re(type-2) of a real grid as float32: the ``.real`` every caller on the path
    keeps (cmtip/alignment/align.py:61-63), from a real fine grid."""
    return _forward(ugrid, H_, K_, L_, support, True, eps, real_out=True)
