"""Differentiable slicing: the type-2 slice of a real autocorrelation as a
``torch.autograd.Function``.

Forward (compute_slices, cmtip/alignment/align.py:39-65):
    I_j = Re sum_k f_k exp(-i k.x_j)
Backward, for an upstream gradient g (real, one value per point):
    dL/df_k     = sum_j g_j cos(k.x_j) = Re(type 1 of g)_k          (the adjoint spread)
    dL/dx_{j,d} = g_j Im(sum_k k_d f_k exp(-i k.x_j))               (an ik-weighted type 2, d = 0, 1, 2)
Both run on the forward's type-2 plan and its point binning: the spread as the plan's
adjoint (CMTIP_EXEC_ADJOINT), the coordinate gradient as type-2 imaginary-part outputs
(real interpolation of the fine grid's imaginary part).  ``rotated_coords`` builds the
rotated pixel coordinates from quaternions with torch operations, so gradients reach
the quaternions for orientation refinement.  ``slice_quaternions`` does the same in one
function without coordinate arrays: the plan bins the rotated pixels straight from the
normalised quaternions (cmtip_nufft_setpts_rotated) and the backward contracts the
coordinate gradient with the pixel positions (one thin matrix product per axis) and
the closed-form derivative of R(q).
"""
import numpy as np
import torch

from ._lib import CMTIP_EXEC_ADJOINT, CMTIP_GRID_REAL, CMTIP_PTS_IMAG, CMTIP_PTS_REAL
from .nufft.plan import DEFAULT_EPS, get_plan


def _modes(M, device):
    return torch.arange(M, device=device, dtype=torch.float32) - (M // 2)


class NufftSlice(torch.autograd.Function):
    """I = Re(T2(f)) at (H, K, L); differentiable in f and in the coordinates."""

    @staticmethod
    def forward(ctx, ac, H, K, L, eps=DEFAULT_EPS):
        ac = ac.contiguous().float()
        x, y, z = (t.contiguous().float().reshape(-1) for t in (H, K, L))
        M = ac.shape[0]
        plan = get_plan(2, M, -1, eps)
        plan.setpts(x, y, z)
        out = torch.empty(x.numel(), dtype=torch.float32, device=ac.device)
        plan.execute(out, ac, CMTIP_GRID_REAL | CMTIP_PTS_REAL)
        ctx.save_for_backward(ac, x, y, z)
        ctx.eps = eps
        ctx.shape = H.shape
        return out.reshape(H.shape)

    @staticmethod
    def backward(ctx, grad):
        ac, x, y, z = ctx.saved_tensors
        eps = ctx.eps
        M = ac.shape[0]
        g = grad.contiguous().float().reshape(-1)
        grad_ac = grad_H = grad_K = grad_L = None
        p2 = get_plan(2, M, -1, eps)
        p2.setpts(x, y, z)  # the forward's points: already registered unless the plan was used since
        if ctx.needs_input_grad[0]:
            # the adjoint of the forward plan: a type-1 spread over the same bins
            grad_ac = torch.empty_like(ac)
            p2.execute(g, grad_ac, CMTIP_GRID_REAL | CMTIP_PTS_REAL | CMTIP_EXEC_ADJOINT, None, 1.0)
        if any(ctx.needs_input_grad[1:4]):
            k = _modes(M, ac.device)
            outs = []
            for d in range(3):
                if not ctx.needs_input_grad[1 + d]:
                    outs.append(None)
                    continue
                shape = [1, 1, 1]
                shape[d] = M
                kf = (ac * k.reshape(shape)).contiguous()
                c = torch.empty(x.numel(), dtype=torch.float32, device=ac.device)
                p2.execute(c, kf, CMTIP_GRID_REAL | CMTIP_PTS_IMAG)  # Im(T2(k_d f))
                outs.append((g * c).reshape(ctx.shape))
            grad_H, grad_K, grad_L = outs
        return grad_ac, grad_H, grad_K, grad_L, None


def nufft_slice(ac, H, K, L, eps=DEFAULT_EPS):
    return NufftSlice.apply(ac, H, K, L, eps)


def quaternion_rotation(q):
    """Rotation matrices of (normalised) scalar-first quaternions, cmtip/eval/geom.py:37-45."""
    q = q / torch.linalg.vector_norm(q, dim=-1, keepdim=True)
    q0, q1, q2, q3 = q.unbind(-1)
    R = torch.stack([
        1 - 2 * (q2 * q2 + q3 * q3), 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2),
        2 * (q1 * q2 + q0 * q3), 1 - 2 * (q1 * q1 + q3 * q3), 2 * (q2 * q3 - q0 * q1),
        2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 1 - 2 * (q1 * q1 + q2 * q2),
    ], -1)
    return R.reshape(q.shape[:-1] + (3, 3))


def rotated_coords(q, pixel_position_reciprocal, reciprocal_extent):
    """x = pi R(q)^-1 s / reciprocal_extent for every (orientation, pixel), differentiable
    in q (R^-1 = R^T for the normalised quaternion).  Returns (H, K, L), each (n, P)."""
    s = torch.as_tensor(pixel_position_reciprocal, device=q.device, dtype=q.dtype).reshape(3, -1)
    Rinv = quaternion_rotation(q).transpose(-1, -2)
    v = torch.einsum("nij,jp->inp", Rinv, s) * (np.pi / float(reciprocal_extent))
    return v[0], v[1], v[2]


# dR_ij / dq_m = sum_l _DR[i, j, m, l] q_l for R = quaternion_rotation(q) of a unit q
# (the partial derivatives of its nine quadratic entries)
_DR_ENTRIES = (
    (0, 0, 2, 2, -4), (0, 0, 3, 3, -4),
    (0, 1, 0, 3, -2), (0, 1, 1, 2, 2), (0, 1, 2, 1, 2), (0, 1, 3, 0, -2),
    (0, 2, 0, 2, 2), (0, 2, 1, 3, 2), (0, 2, 2, 0, 2), (0, 2, 3, 1, 2),
    (1, 0, 0, 3, 2), (1, 0, 1, 2, 2), (1, 0, 2, 1, 2), (1, 0, 3, 0, 2),
    (1, 1, 1, 1, -4), (1, 1, 3, 3, -4),
    (1, 2, 0, 1, -2), (1, 2, 1, 0, -2), (1, 2, 2, 3, 2), (1, 2, 3, 2, 2),
    (2, 0, 0, 2, -2), (2, 0, 1, 3, 2), (2, 0, 2, 0, -2), (2, 0, 3, 1, 2),
    (2, 1, 0, 1, 2), (2, 1, 1, 0, 2), (2, 1, 2, 3, 2), (2, 1, 3, 2, 2),
    (2, 2, 1, 1, -4), (2, 2, 2, 2, -4),
)
_DR_CACHE = {}


def rotation_derivative(device, dtype):
    """(3, 3, 4, 4) coefficients of dR/dq (see _DR_ENTRIES)."""
    key = (str(device), dtype)
    D = _DR_CACHE.get(key)
    if D is None:
        D = torch.zeros((3, 3, 4, 4), dtype=dtype)
        for i, j, m, l, v in _DR_ENTRIES:
            D[i, j, m, l] = v
        D = D.to(device)
        _DR_CACHE[key] = D
    return D


class SliceQuaternions(torch.autograd.Function):
    """I = Re(T2(f)) at the rotated pixels x = pi R(q)^T s / ext (q normalised), with no
    coordinate arrays: the forward bins the pixels straight from the quaternions, the
    backward gives d/df (the adjoint spread) and d/dq (the coordinate gradient
    contracted with s, then the derivative of R and of the normalisation)."""

    @staticmethod
    def forward(ctx, ac, q, s64, ext, eps=DEFAULT_EPS):
        ac = ac.contiguous().float()
        norm = torch.linalg.vector_norm(q, dim=-1, keepdim=True)
        qn = (q / norm).contiguous()
        M = ac.shape[0]
        n, P = q.shape[0], s64.shape[1]
        plan = get_plan(2, M, -1, eps)
        plan.setpts_rotated(qn, s64, ext)
        out = torch.empty(n * P, dtype=torch.float32, device=ac.device)
        plan.execute(out, ac, CMTIP_GRID_REAL | CMTIP_PTS_REAL)
        ctx.save_for_backward(ac, qn, norm, s64)
        ctx.ext, ctx.eps = ext, eps
        return out.reshape(n, P)

    @staticmethod
    def backward(ctx, grad):
        ac, qn, norm, s64 = ctx.saved_tensors
        M = ac.shape[0]
        n, P = qn.shape[0], s64.shape[1]
        g = grad.contiguous().float().reshape(-1)
        grad_ac = grad_q = None
        p2 = get_plan(2, M, -1, ctx.eps)
        p2.setpts_rotated(qn, s64, ctx.ext)  # skipped while the forward's points are registered
        if ctx.needs_input_grad[0]:
            grad_ac = torch.empty_like(ac)
            p2.execute(g, grad_ac, CMTIP_GRID_REAL | CMTIP_PTS_REAL | CMTIP_EXEC_ADJOINT, None, 1.0)
        if ctx.needs_input_grad[1]:
            k = _modes(M, ac.device)
            s32 = s64.to(torch.float32)
            # G[:, d, j] = sum_p dL/dx_d(o, p) s_j(p): x_d = c sum_j R[j, d] s_j with c = pi / ext
            G = torch.empty((n, 3, 3), dtype=torch.float32, device=ac.device)
            c = torch.empty(n * P, dtype=torch.float32, device=ac.device)
            for d in range(3):
                shape = [1, 1, 1]
                shape[d] = M
                kf = (ac * k.reshape(shape)).contiguous()
                p2.execute(c, kf, CMTIP_GRID_REAL | CMTIP_PTS_IMAG)  # Im(T2(k_d f))
                G[:, d, :] = (g * c).reshape(n, P) @ s32.t()
            scale = float(np.float32(np.pi)) / float(np.float32(ctx.ext))
            gR = G.transpose(1, 2) * scale  # dL/dR[j, d]
            dqn = torch.einsum("nij,ijml,nl->nm", gR, rotation_derivative(ac.device, torch.float32), qn)
            grad_q = (dqn - qn * (qn * dqn).sum(-1, keepdim=True)) / norm
        return grad_ac, grad_q, None, None, None


def slice_quaternions(ac, q, pixel_position_reciprocal, reciprocal_extent, eps=DEFAULT_EPS):
    """Slices (n, P) of ac at quaternions q, differentiable in ac and q.  Float32
    quaternions on the device take SliceQuaternions (no coordinate arrays); others the
    torch coordinate map and nufft_slice."""
    if q.dtype == torch.float32 and q.is_cuda and reciprocal_extent:
        s64 = torch.as_tensor(pixel_position_reciprocal, device=q.device, dtype=torch.float64).reshape(3, -1)
        ext = float(np.float32(reciprocal_extent.item() if isinstance(reciprocal_extent, torch.Tensor)
                               else reciprocal_extent))
        return SliceQuaternions.apply(ac, q, s64.contiguous(), ext, eps)
    H, K, L = rotated_coords(q, pixel_position_reciprocal, reciprocal_extent)
    return nufft_slice(ac, H, K, L, eps)


def l2_loss(model, images, gamma=None):
    """sum (gamma model - images)^2 (the matching distance, align.py:122-128, unweighted)."""
    if gamma is None:
        return ((model - images) ** 2).sum()
    return ((gamma * model - images) ** 2).sum()


def poisson_nll(model, images, floor=1e-6):
    """sum (model - images log model) for Poisson counts (model clamped at floor)."""
    m = torch.clamp(model, min=floor)
    return (m - images * torch.log(m)).sum()


def refine_orientations(ac, q0, images, pixel_position_reciprocal, reciprocal_extent, steps=50, lr=2e-3,
                        loss="l2", eps=DEFAULT_EPS):
    """Gradient refinement of quaternions q0 (n, 4) against images (n, P); returns the
    refined, normalised quaternions (scalar part >= 0)."""
    q = q0.detach().clone().float().requires_grad_(True)
    opt = torch.optim.Adam([q], lr=lr)
    imgs = images.float()
    ac = ac.float()
    for _ in range(steps):
        opt.zero_grad()
        model = slice_quaternions(ac, q, pixel_position_reciprocal, reciprocal_extent, eps)
        if loss == "l2":
            obj = l2_loss(model, imgs) / imgs.numel()
        else:
            obj = poisson_nll(model, imgs) / imgs.numel()
        obj.backward()
        opt.step()
    with torch.no_grad():
        qn = q / torch.linalg.vector_norm(q, dim=-1, keepdim=True)
        qn = torch.where(qn[:, :1] < 0, -qn, qn)
    return qn.detach()
