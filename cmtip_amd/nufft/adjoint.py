"""Adjoint (type 1) NUFFT: nonuniform points -> grid, isign +1, divided by M^3.

Mirrors cmtip/nufft/adjoint.py:13-86.  Both names run the HIP path.
"""
import numpy as np
import torch

from .. import _arrays
from .._lib import CMTIP_GRID_REAL, CMTIP_PTS_BOUND, CMTIP_PTS_ONES, CMTIP_PTS_REAL
from .plan import DEFAULT_EPS, get_plan, tuned_setpts


def _adjoint(nuvect, H_, K_, L_, M, use_recip_sym, support, eps, scale=None, ones=False, autotune=False,
             rot=None, check_range=True):
    """rot (quats, s, recip_extent), device tensors: the points are the rotated pixels of
    gen_nonuniform_positions(quats, s, recip_extent=...), registered without coordinate
    arrays (H_, K_, L_ unused; no autotune).  check_range False leaves the range check of
    adjoint.py:29,66 to the caller (``assert_points_in_range``), so that several spreads
    can be queued before one synchronising read; out-of-range points are clamped by the
    kernels, never read or written outside the grid."""
    if rot is None:
        as_torch = _arrays.is_torch(nuvect, H_)
        x = _arrays.to_dev(H_, torch.float32).reshape(-1)
        y = _arrays.to_dev(K_, torch.float32).reshape(-1)
        z = _arrays.to_dev(L_, torch.float32).reshape(-1)
        npts, dev = x.numel(), x.device
    else:
        assert not autotune, "autotuning bins registered coordinate arrays again"
        as_torch = True
        quats, s, ext = rot
        npts, dev = quats.shape[0] * s.shape[1], quats.device
    flags = 0
    c = None
    if ones:
        flags |= CMTIP_PTS_ONES
    else:
        cplx = torch.is_complex(nuvect) if isinstance(nuvect, torch.Tensor) else np.iscomplexobj(nuvect)
        if cplx:
            c = _arrays.to_dev(nuvect, torch.complex64).reshape(-1)
        else:
            c = _arrays.to_dev(nuvect, torch.float32).reshape(-1)
            flags |= CMTIP_PTS_REAL
        assert c.numel() == npts
    M = int(M)
    if use_recip_sym:
        f = torch.empty((M, M, M), dtype=torch.float32, device=dev)
        flags |= CMTIP_GRID_REAL
    else:
        f = torch.empty((M, M, M), dtype=torch.complex64, device=dev)
    sup = None if support is None else _arrays.to_dev(support, torch.float32)
    plan = get_plan(1, M, +1, eps)
    scale = float(1.0 / M ** 3) if scale is None else float(scale)
    bind = not autotune and c is not None and (flags & CMTIP_PTS_REAL)
    if rot is not None:
        plan.setpts_rotated(quats, s, ext, c if bind else None)
        if bind:
            flags |= CMTIP_PTS_BOUND
            c = None
    elif autotune:
        tuned_setpts(plan, x, y, z, c, f, flags, sup, scale)
    elif bind:
        # real strengths ride in the bin-sorted point records: no gather of c by index
        plan.setpts_strengths(x, y, z, c)
        flags |= CMTIP_PTS_BOUND
        c = None
    else:
        plan.setpts(x, y, z)
    # the range check of adjoint.py:29, 66: the maximum is recorded while setpts bins
    if check_range:
        assert_points_in_range(plan)
    plan.execute(c, f, flags, sup, scale)
    return _arrays.out(f, as_torch)


def assert_points_in_range(*plans):
    """The reference's assertion max |coordinate| < 3 pi (adjoint.py:29,66) on the points
    registered with each plan (one stream synchronisation for all of them)."""
    for plan in plans:
        assert plan.points_max_abs() < 3 * np.pi


def adjoint_cpu(nuvect, H_, K_, L_, M, use_recip_sym=True, support=None, *, eps=DEFAULT_EPS):
    """Type-1 NUFFT: f_k = sum_j c_j exp(+i k.x_j), times support, real part when
    use_recip_sym, divided by M^3.  Signature of cmtip/nufft/adjoint.py:13; runs on the
    GPU (the finufft call at adjoint.py:33 is replaced)."""
    return _adjoint(nuvect, H_, K_, L_, M, use_recip_sym, support, eps)


def adjoint_gpu(nuvect, H_, K_, L_, M, use_recip_sym=True, support=None, *, eps=DEFAULT_EPS):
    """Type-1 NUFFT; signature of cmtip/nufft/adjoint.py:46."""
    return _adjoint(nuvect, H_, K_, L_, M, use_recip_sym, support, eps)


def adjoint_ones(H_, K_, L_, M, *, scale=None, eps=DEFAULT_EPS):
    """Real type-1 of unit strengths (the kernel of autocorrelation.py:143-147) without
    materialising the ones; scale defaults to 1/M^3 as adjoint_* divides."""
    return _adjoint(None, H_, K_, L_, M, True, None, eps, scale=scale, ones=True)
