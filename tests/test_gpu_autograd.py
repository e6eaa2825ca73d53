"""Differentiable slicing (cmtip_amd.autograd) against FP64 finite differences of the
exact type-2 sum and against the exact type-1 adjoint."""
import numpy as np
import pytest
import torch

from cmtip_amd import autograd as AG
from cmtip_amd import synth
from oracle import nufft as O
from tests.gpu_util import dev, rel_l2, say

pytestmark = pytest.mark.gpu


def _setup(M=21, n_orient=64, side=32, seed=0):
    d = dev()
    s, ext = synth.detector(side)
    ac = synth.gaussian_ac(M, 30, seed=seed)
    q = synth.random_quaternions(n_orient, seed=seed + 1)
    return d, s, ext, ac, q


def _coords64(q, s, ext):
    """x = pi R(q)^T s / ext in float64 (the differentiable map of AG.rotated_coords)."""
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    q0, q1, q2, q3 = q.T
    R = np.stack([1 - 2 * (q2 * q2 + q3 * q3), 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2),
                  2 * (q1 * q2 + q0 * q3), 1 - 2 * (q1 * q1 + q3 * q3), 2 * (q2 * q3 - q0 * q1),
                  2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 1 - 2 * (q1 * q1 + q2 * q2)], -1)
    R = R.reshape(-1, 3, 3)
    v = np.einsum("nji,jp->inp", R, s.reshape(3, -1).astype(np.float64)) * (np.pi / float(ext))
    return v[0].ravel(), v[1].ravel(), v[2].ravel()


def test_grad_ac_is_exact_adjoint():
    d, s, ext, ac, q = _setup()
    x, y, z = _coords64(q, s, ext)
    rng = np.random.default_rng(0)
    g = rng.standard_normal(x.size)
    f = torch.from_numpy(ac).to(d).requires_grad_(True)
    X, Y, Z = (torch.from_numpy(v.astype(np.float32)).to(d) for v in (x, y, z))
    say("autograd grad_ac vs exact adjoint")
    I = AG.nufft_slice(f, X, Y, Z)
    (I * torch.from_numpy(g.astype(np.float32)).to(d)).sum().backward()
    want = O.nudft1(g, X.cpu().numpy(), Y.cpu().numpy(), Z.cpu().numpy(), ac.shape[0]).real
    assert rel_l2(f.grad, want) < 1e-5
    # linearity: <grad, f> equals the loss
    L = float((I.detach().double() * torch.from_numpy(g).to(d)).sum())
    assert abs(float((f.grad.double() * f.detach().double()).sum()) - L) < 1e-4 * abs(L)


def test_grad_coords_vs_fp64_central_differences():
    d, s, ext, ac, q = _setup()
    x, y, z = _coords64(q, s, ext)
    x, y, z = (v.astype(np.float32).astype(np.float64) for v in (x, y, z))
    rng = np.random.default_rng(1)
    g = rng.standard_normal(x.size)
    v = rng.standard_normal((3, x.size))
    h = 1e-6

    def loss(dx):
        c = O.nufft2(ac.astype(np.float64), x + dx[0], y + dx[1], z + dx[2], eps=1e-13).real
        return float(np.dot(g, c))
    say("autograd coordinate gradient vs FP64 central differences")
    fd = (loss(h * v) - loss(-h * v)) / (2 * h)
    X, Y, Z = (torch.from_numpy(t.astype(np.float32)).to(d).requires_grad_(True) for t in (x, y, z))
    f = torch.from_numpy(ac).to(d)
    (AG.nufft_slice(f, X, Y, Z) * torch.from_numpy(g.astype(np.float32)).to(d)).sum().backward()
    an = sum(float((t.grad.double().cpu() * torch.from_numpy(v[i])).sum()) for i, t in enumerate((X, Y, Z)))
    assert abs(an - fd) < 1e-4 * abs(fd)
    # per-axis check too
    for i, t in enumerate((X, Y, Z)):
        e = np.zeros((3, x.size))
        e[i] = v[i]
        fdi = (loss(h * e) - loss(-h * e)) / (2 * h)
        assert abs(float((t.grad.double().cpu() * torch.from_numpy(v[i])).sum()) - fdi) < 1e-4 * abs(fdi)


def test_grad_quaternions_vs_fp64_central_differences():
    d, s, ext, ac, q = _setup(n_orient=16)
    rng = np.random.default_rng(2)
    g = rng.standard_normal(16 * s.shape[-1])
    v = rng.standard_normal(q.shape)
    h = 1e-6

    def loss(qq):
        x, y, z = _coords64(qq, s, ext)
        return float(np.dot(g, O.nufft2(ac.astype(np.float64), x, y, z, eps=1e-13).real))
    q64 = q.astype(np.float64)
    fd = (loss(q64 + h * v) - loss(q64 - h * v)) / (2 * h)
    say("autograd quaternion gradient vs FP64 central differences")
    qt = torch.from_numpy(q).to(d).requires_grad_(True)
    I = AG.slice_quaternions(torch.from_numpy(ac).to(d), qt, s, ext)
    (I.reshape(-1) * torch.from_numpy(g.astype(np.float32)).to(d)).sum().backward()
    an = float((qt.grad.double().cpu() * torch.from_numpy(v)).sum())
    assert abs(an - fd) < 2e-4 * max(abs(fd), 1e-12)


@pytest.mark.parametrize("loss", ["l2", "poisson"])
def test_refinement_from_two_degrees(loss):
    d = dev()
    M, side, n = 41, 32, 8
    s, ext = synth.detector(side)
    ac = torch.from_numpy(synth.gaussian_ac(M, 100, seed=4)).to(d)
    q_true = torch.from_numpy(synth.random_quaternions(n, seed=5)).to(d)
    with torch.no_grad():
        clean = AG.slice_quaternions(ac, q_true, s, ext)
        scale = 10.0 / float(clean.clamp(min=0).mean())
        imgs = clean * scale if loss == "l2" else clean.clamp(min=0) * scale
    # perturb by 2 degrees about random axes
    rng = np.random.default_rng(6)
    axis = rng.standard_normal((n, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    th = np.deg2rad(2.0)
    dq = np.concatenate([np.full((n, 1), np.cos(th / 2)), np.sin(th / 2) * axis], 1)
    qt = q_true.cpu().numpy().astype(np.float64)
    # Hamilton product q_true * dq
    w1, x1, y1, z1 = qt.T
    w2, x2, y2, z2 = dq.T
    q0 = np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                   w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], 1)
    from cmtip_amd.alignment import errors
    before = errors.alignment_error(q0, qt)
    assert np.allclose(before, 2.0, atol=0.01)
    say(f"refinement {loss}")
    qr = AG.refine_orientations(ac * scale, torch.from_numpy(q0.astype(np.float32)).to(d), imgs, s, ext,
                                steps=80, lr=2e-3, loss=loss)
    after = errors.alignment_error(qr.cpu().numpy().astype(np.float64), qt)
    assert np.all(after < 0.5), after


@pytest.mark.parametrize("n,side", [(3, 17), (40, 32)])
def test_slice_quaternions_matches_coordinate_path(n, side):
    # SliceQuaternions (no coordinate arrays, closed-form dR/dq) against the torch
    # coordinate map + NufftSlice on the same inputs: slices, d/dac and d/dq
    d = dev()
    M = 21
    s, ext = synth.detector(side)
    ac = torch.from_numpy(synth.gaussian_ac(M, 30, seed=7)).to(d)
    q0 = torch.from_numpy(synth.random_quaternions(n, seed=8)).to(d) * 1.3  # not normalised
    rng = np.random.default_rng(9)
    g = torch.from_numpy(rng.standard_normal((n, side * side)).astype(np.float32)).to(d)
    say(f"slice_quaternions vs coordinate path n={n} side={side}")
    outs = []
    for fused in (True, False):
        q = q0.clone().requires_grad_(True)
        f = ac.clone().requires_grad_(True)
        if fused:
            I = AG.slice_quaternions(f, q, s, ext)
        else:
            I = AG.nufft_slice(f, *AG.rotated_coords(q, s, ext))
        (I * g).sum().backward()
        outs.append((I.detach(), f.grad, q.grad))
    (I1, fa1, qa1), (I2, fa2, qa2) = outs
    assert rel_l2(I1, I2) < 1e-5
    assert rel_l2(fa1, fa2) < 1e-5
    assert rel_l2(qa1, qa2) < 1e-4
