"""Parity at the sizes of the larger workloads: BF16 matching with top-8 FP32 rescoring
at 20000 images x 50000 references x 16384 pixels, masked matching on an M = 151 grid
in two detector binnings, and the M = 151 autocorrelation solve with autotuned spreads.

The FP64 references are computed on the GPU in float64 (torch) from the same float32
model slices (the slices themselves are pinned against the oracle in
test_gpu_slices.py); assignments must be identical for every image whose best-to-
second-best squared distance gap exceeds 1e-5 of the best."""
import numpy as np
import pytest
import torch

from cmtip_amd import alignment, synth
from tests.gpu_util import dev, rel_l2, say

pytestmark = pytest.mark.gpu


def _fp64_best(A, B, w2, gamma, mask=None, chunk=2000):
    """Per image: (argmin, best, second best) of d2 = sum_p [m] w2 (gamma a - b)^2 in FP64,
    images in chunks (A: float32 (n_ref, P) on the device)."""
    n_img = B.shape[0]
    best_i = torch.empty(n_img, dtype=torch.int64, device=B.device)
    d01 = torch.empty((2, n_img), dtype=torch.float64, device=B.device)
    A64 = A.double()
    w = w2.double()
    an = ((gamma * A64) ** 2 * w).sum(1) if mask is None else None
    for i0 in range(0, n_img, chunk):
        b = B[i0:i0 + chunk].double()
        if mask is None:
            D = an[:, None] - 2 * gamma * (A64 * w) @ b.T + ((b * b) * w).sum(1)[None, :]
        else:
            mw = mask[i0:i0 + chunk].double() * w
            D = ((gamma * A64) ** 2) @ mw.T - 2 * gamma * A64 @ (mw * b).T + (mw * b * b).sum(1)[None, :]
        v, idx = torch.topk(D, 2, dim=0, largest=False)
        best_i[i0:i0 + chunk] = idx[0]
        d01[:, i0:i0 + chunk] = v
        del D
    return best_i, d01


def test_bf16_top8_at_full_size():
    # 20000 images x 50000 references x 16384 pixels (M = 81), Poisson counts at 5 photons
    # per pixel, log1p-scaled; BF16 cross term (library GEMM in 1 GiB blocks), top-8
    # rescored in FP32
    d = dev()
    M, side, n_ref, n_img = 81, 128, 50000, 20000
    s, ext = synth.detector(side)
    ac = torch.from_numpy(synth.gaussian_ac(M, 1000, seed=0)).to(d)
    ref = synth.random_quaternions(n_ref, seed=0)
    q_img = synth.random_quaternions(n_img, seed=1)
    sl = alignment.Slicer(ac, s, ext)
    say(f"full-size bf16 matching {n_img} x {n_ref} x {side * side}")
    clean = sl.slice(q_img).clamp(min=0)
    torch.manual_seed(2)
    imgs = torch.log1p(torch.poisson(clean * (5.0 / float(clean.mean()))))
    del clean
    q, idx = alignment.match_orientations(0, s, ext, imgs, ac, n_ref, ref_orientations=ref, precision="bf16",
                                         return_index=True)
    say("  the same with this library's BF16 kernel (cross='fused')")
    qf, idxf = alignment.match_orientations(0, s, ext, imgs, ac, n_ref, ref_orientations=ref, precision="bf16",
                                           return_index=True, cross="fused")
    A = sl.slice(torch.from_numpy(ref).to(d))
    P = side * side
    w2 = torch.full((P,), 1.0 / P ** 2, device=d)
    gamma = float(imgs.double().std(unbiased=False) / A.double().std(unbiased=False))
    want, d01 = _fp64_best(A, imgs, w2, gamma)
    ok = ((d01[1] - d01[0]) > 1e-5 * d01[0].abs()).cpu().numpy()
    got = idx.cpu().numpy()
    w = want.cpu().numpy()
    agree = float((got == w).mean())
    say(f"agreement with FP64 argmin: {agree:.5f} overall, images outside the tie band: {ok.mean():.4f}")
    assert ok.mean() > 0.9
    assert np.array_equal(got[ok], w[ok]), f"{(got[ok] != w[ok]).sum()} mismatches of {ok.sum()}"
    gf = idxf.cpu().numpy()
    assert np.array_equal(gf[ok], w[ok]), f"fused: {(gf[ok] != w[ok]).sum()} mismatches of {ok.sum()}"


@pytest.mark.parametrize("side", [128, 96])
def test_masked_matching_m151(side):
    # M = 151 grid (5000-atom density), per-image beamstop masks of uneven size.  Masking
    # the bright centre shrinks the best-to-second gaps ~100 x (median 5.6e-4 relative at
    # 128 x 128); with y^2 split into two bfloat16 parts the BF16 top-8 still holds the
    # best reference (a single bfloat16 y^2 missed it for 7 % of the images), so both
    # forms must give the FP64 assignment outside the tie band; 10000 references x 10000
    # images per binning
    d = dev()
    M, n_ref, n_img = 151, 10000, 10000
    s, ext = synth.detector(side)
    ac = torch.from_numpy(synth.gaussian_ac(M, 5000, seed=0)).to(d)
    ref = synth.random_quaternions(n_ref, seed=3)
    q_img = synth.random_quaternions(n_img, seed=4)
    sl = alignment.Slicer(ac, s, ext)
    say(f"masked matching M=151 side={side} {n_img} x {n_ref}")
    clean = sl.slice(q_img).clamp(min=0)
    torch.manual_seed(5)
    imgs = torch.poisson(clean * (10.0 / float(clean.mean())))
    P = side * side
    rng = np.random.default_rng(side)
    yy, xx = np.mgrid[:side, :side]
    masks = np.ones((n_img, P), np.uint8)
    for i in range(n_img):
        cy, cx = rng.integers(side // 3, 2 * side // 3, 2)
        r = rng.uniform(2, side / 5)
        masks[i] = (((yy - cy) ** 2 + (xx - cx) ** 2) > r * r).ravel()
    mask_d = torch.from_numpy(masks).to(d)
    q, idx = alignment.match_orientations(0, s, ext, imgs, ac, n_ref, ref_orientations=ref, precision="fp32",
                                         return_index=True, masks=masks)
    q16, idx16 = alignment.match_orientations(0, s, ext, imgs, ac, n_ref, ref_orientations=ref,
                                             precision="bf16", return_index=True, masks=masks)
    say("  bf16 with this library's kernel (cross='fused', 16 candidates)")
    q16f, idx16f = alignment.match_orientations(0, s, ext, imgs, ac, n_ref, ref_orientations=ref,
                                               precision="bf16", return_index=True, masks=masks, cross="fused")
    A = sl.slice(torch.from_numpy(ref).to(d))
    w2 = torch.full((P,), 1.0 / P ** 2, device=d)
    valid = imgs[mask_d.bool()].double()
    gamma = float(valid.std(unbiased=False) / A.double().std(unbiased=False))
    want, d01 = _fp64_best(A, imgs, w2, gamma, mask=mask_d)
    ok = ((d01[1] - d01[0]) > 1e-5 * d01[0].abs()).cpu().numpy()
    got, w = idx.cpu().numpy(), want.cpu().numpy()
    say(f"  outside the tie band {ok.mean():.4f}; fp32 agreement {(got == w).mean():.5f}, "
        f"bf16 {(idx16.cpu().numpy() == w).mean():.5f}")
    assert ok.mean() > 0.9
    assert np.array_equal(got[ok], w[ok]), f"{(got[ok] != w[ok]).sum()} mismatches of {ok.sum()}"
    got16 = idx16.cpu().numpy()
    bad = np.flatnonzero(ok & (got16 != w))
    for i in bad[:5]:
        # FP64 distance of the BF16 choice against the best and second best
        mw = mask_d[i].double() * w2.double()
        dd = ((gamma * A[[got16[i], w[i]]].double() - imgs[i].double()) ** 2 * mw).sum(1)
        say(f"  bf16 mismatch image {i}: d2(bf16 pick) {float(dd[0]):.9e}, d2(best) {float(dd[1]):.9e}, "
            f"second {float(d01[1, i]):.9e}, relative gap {float((dd[0] - dd[1]) / dd[1]):.2e}")
    assert np.array_equal(got16[ok], w[ok]), f"bf16: {(got16[ok] != w[ok]).sum()} mismatches of {ok.sum()}"
    g16f = idx16f.cpu().numpy()
    assert np.array_equal(g16f[ok], w[ok]), f"bf16 fused: {(g16f[ok] != w[ok]).sum()} mismatches of {ok.sum()}"


def test_solve_m151_autotuned():
    # M = 151 solve over 1000 images of 96 x 96 pixels (9.2 M points): spreads with
    # autotuned bin settings give the same solution as the defaults, bit for bit
    from cmtip_amd import nufft
    from cmtip_amd.autocorrelation import solve_ac
    d = dev()
    M, side, n_img = 151, 96, 1000
    s, ext = synth.detector(side)
    ac = torch.from_numpy(synth.gaussian_ac(M, 5000, seed=0)).to(d)
    q = torch.from_numpy(synth.random_quaternions(n_img, seed=6)).to(d)
    imgs = alignment.Slicer(ac, s, ext).slice(q).clamp(min=0)
    say("solve M=151 default / autotuned")
    a0, i0 = solve_ac(0, s, ext, imgs.reshape(n_img, 1, -1), M, orientations=q, maxiter=10, rlambda=1e-4,
                      return_info=True)
    a1, i1 = solve_ac(0, s, ext, imgs.reshape(n_img, 1, -1), M, orientations=q, maxiter=10, rlambda=1e-4,
                      return_info=True, autotune=True)
    assert i0["iterations"] == i1["iterations"] == 10
    # real spreads sum exact 64-bit fixed-point contributions, so other bin settings give
    # the same spreads bit for bit, and the CG (fixed-order FP64 reductions) the same solve
    say(f"  rel L2 autotuned vs default {rel_l2(a1, a0):.3e}")
    assert torch.equal(a1, a0)
    nufft.clear_plans()
