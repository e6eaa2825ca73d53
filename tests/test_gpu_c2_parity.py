"""Parity at the production sizes: the M = 81 workload (128 x 128-pixel images, the
q-vectors of 5000 image orientations, 81.92 M points) and the M = 41 matching workload
(1000 references x 256 Poisson images of 64 x 64 pixels), against the FP64 oracle.

At this size the spreading and interpolation kernels run in the regime the benchmark
times: every image's central pixels fall into the bins around the origin, which then
hold many point chunks each, and the 162-mode kernel grid is 324^3.  Exact sums are
evaluated on samples: type 2 at selected points (full M^3 sum per point), type 1 at a
tensor product of selected modes (oracle.nufft.nudft1_modes, linear in the points).
"""
import time

import numpy as np
import pytest
import torch

from cmtip_amd import alignment, autocorrelation, nufft, synth
from cmtip_amd.autocorrelation import autocorrelation as AC
from oracle import ac as OAC
from oracle import align as OA
from oracle import geom
from oracle import nufft as O
from tests.gpu_util import dev, rel_l2, say

pytestmark = pytest.mark.gpu

M81, SIDE, N_IMG = 81, 128, 5000


@pytest.fixture(scope="module")
def c2():
    """Image q-vectors of the M = 81 workload (5000 orientations x 16384 pixels), the
    autocorrelation and Poisson images (10 photons per pixel) on the device."""
    d = dev()
    s, ext = synth.detector(SIDE)
    ac = synth.gaussian_ac(M81, 1000, seed=0)
    q = synth.random_quaternions(N_IMG, seed=1)
    say(f"c2 inputs: M={M81} images={N_IMG} pixels={SIDE * SIDE}")
    x, y, z = AC.gen_nonuniform_positions(torch.from_numpy(q).to(d), torch.from_numpy(s).to(d), recip_extent=ext)
    x, y, z = (t.reshape(-1).contiguous() for t in (x, y, z))
    ac_d = torch.from_numpy(ac).to(d)
    clean = alignment.Slicer(ac_d, s, ext).slice(torch.from_numpy(q).to(d)).clamp_(min=0)
    torch.manual_seed(2)
    imgs = torch.poisson(clean * (10.0 / float(clean.mean()))).reshape(-1)
    del clean
    torch.cuda.synchronize()
    return {"s": s, "ext": ext, "ac": ac, "ac_d": ac_d, "q": q, "x": x, "y": y, "z": z, "data": imgs,
            "hx": x.cpu().numpy(), "hy": y.cpu().numpy(), "hz": z.cpu().numpy()}


def _sample_points(hx, hy, hz, n_rand=192, n_center=64, seed=5):
    """Random points plus the points nearest the origin (the densest bins)."""
    rng = np.random.default_rng(seed)
    r2 = hx.astype(np.float64) ** 2 + hy.astype(np.float64) ** 2 + hz.astype(np.float64) ** 2
    centre = np.argpartition(r2, n_center)[:n_center]
    return np.unique(np.concatenate([rng.choice(hx.size, n_rand, replace=False), centre]))


def _sample_modes(M, seed):
    """Six mode numbers per axis: 0, +-1, both ends of the range and one at random."""
    lo, hi = -(M // 2), M - M // 2 - 1
    r = int(np.random.default_rng(seed).integers(lo + 2, hi - 1))
    k = np.unique(np.array([lo, -1, 0, 1, r, hi]))
    return k


def test_type2_c2_image_points(c2):
    # forward.py:13-42 on 81.92 M points at M = 81 (real grid, the .real callers keep)
    say("type 2 at M=81 over 81.92 M image points")
    got = nufft.forward_gpu(c2["ac_d"], c2["x"], c2["y"], c2["z"])
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    idx = _sample_points(c2["hx"], c2["hy"], c2["hz"])
    t = time.time()
    want = O.nudft2(c2["ac"].astype(np.float64), c2["hx"][idx], c2["hy"][idx], c2["hz"][idx])
    say(f"  exact sums at {idx.size} points in {time.time() - t:.1f}s")
    err = rel_l2(got[idx], want)
    say(f"  type 2 rel L2 {err:.2e}")
    assert err < 1e-5
    # the real-output form (what slicing keeps) agrees with the real part
    got_r = nufft.forward_real(c2["ac_d"], c2["x"], c2["y"], c2["z"]).cpu().numpy()
    assert rel_l2(got_r[idx], want.real) < 1e-5


@pytest.mark.parametrize("modes,ones", [(81, False), (162, True)])
def test_type1_c2_image_points(c2, modes, ones):
    # adjoint.py:13-43 over 81.92 M points: the rhs spread (Poisson image strengths, 81
    # modes) and the PSF kernel spread (unit strengths, 162 modes) of
    # autocorrelation.py:141-168, unnormalised (scale 1) for the comparison
    say(f"type 1 at {modes} modes over 81.92 M points, {'unit' if ones else 'image'} strengths")
    x, y, z = c2["x"], c2["y"], c2["z"]
    if ones:
        got = nufft.adjoint_ones(x, y, z, modes, scale=1.0)
        c = np.ones(x.numel())
    else:
        got = nufft.adjoint_gpu(c2["data"], x, y, z, modes) * float(modes) ** 3
        c = c2["data"].cpu().numpy().astype(np.float64)
    torch.cuda.synchronize()
    got = got.cpu().numpy().astype(np.float64)
    k = _sample_modes(modes, seed=modes)
    t = time.time()
    want = O.nudft1_modes(c, c2["hx"], c2["hy"], c2["hz"], k, k, k).real
    say(f"  exact sums at {k.size ** 3} modes in {time.time() - t:.1f}s")
    i = k + modes // 2
    sub = got[np.ix_(i, i, i)]
    err = rel_l2(sub, want)
    # away from the zero mode (which holds sum(c) and dominates the norm)
    nz = np.ones(want.shape, bool)
    nz[tuple(np.searchsorted(k, 0) for _ in range(3))] = False
    err_nz = rel_l2(sub[nz], want[nz])
    say(f"  type 1 rel L2 {err:.2e} (without the zero mode {err_nz:.2e})")
    assert err < 1e-5
    assert err_nz < 1e-5


def test_solve_ac_m81_vs_oracle():
    # solve_ac (autocorrelation.py:176-248) at M = 81 on 1000 images of 128 x 128 pixels
    # (16.4 M points), 20 CG iterations, Tikhonov weight 1e-4, against the oracle's FP64
    # operators (ES NUFFT at the same eps) and scipy's cg
    M, n_img = M81, 1000
    s, ext = synth.detector(SIDE)
    ac = synth.gaussian_ac(M, 1000, seed=0)
    q = synth.random_quaternions(n_img, seed=1)
    d = dev()
    imgs = alignment.Slicer(torch.from_numpy(ac).to(d), s, ext).slice(torch.from_numpy(q).to(d))
    imgs = imgs.cpu().numpy().reshape(n_img, 1, -1)
    say(f"solve_ac M={M} images={n_img} 20 iterations rho=1e-4")
    got, info = autocorrelation.solve_ac(0, s, ext, imgs, M, orientations=q, maxiter=20, rtol=0.0, rlambda=1e-4,
                                         return_info=True)
    t = time.time()
    want, iters, sup, dvec, F = OAC.solve(q, s, ext, imgs, M, maxiter=20, rlambda=1e-4, eps=1e-6, rtol=0.0)
    say(f"  oracle solve in {time.time() - t:.1f}s")
    assert info["iterations"] == iters == 20
    err = rel_l2(got, want)
    # residual of the oracle's FP64 system at both solutions
    anti = OAC.gen_F_antisupport(M)

    def W(u):
        return (OAC.core_problem_convolution(u, M, F, F.shape[0], sup) + 1e-4 * u
                + 1e3 * OAC.fourier_reg(u, sup, anti, M))
    r_got = np.linalg.norm(W(np.asarray(got, np.float64).ravel()) - dvec) / np.linalg.norm(dvec)
    r_want = np.linalg.norm(W(want.ravel()) - dvec) / np.linalg.norm(dvec)
    say(f"  rel L2 {err:.2e}; residual {r_got:.3e} (oracle {r_want:.3e})")
    # FP32 CG (FP64 scalars) against FP64 CG after 20 iterations: measured 1.4e-4 rel L2,
    # and a residual of the oracle's system 1.37x the oracle's own (the FP32 operator's
    # kernel spectrum differs from the FP64 one by ~1e-7 relative, amplified by W's
    # conditioning); bounds 5e-4 and 1.5x
    assert err < 5e-4
    assert r_got < 1.5 * r_want


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_c1_end_to_end(precision):
    # the M = 41 matching workload in full: 200-atom autocorrelation, 64 x 64 detector,
    # 1000 references (seed 0), 256 images = FP64 slices at random orientations (seed 1)
    # with Poisson noise at 10 photons per pixel (seed 2); match_orientations
    # (align.py:68-132) against sklearn euclidean_distances + argmin on FP64 slices
    M, side, n_ref, n_img = 41, 64, 1000, 256
    s, ext = synth.detector(side)
    ac = synth.gaussian_ac(M, 200, seed=0, dtype=np.float64)
    ref = synth.random_quaternions(n_ref, seed=0)
    q_img = synth.random_quaternions(n_img, seed=1)
    x, y, z = geom.rotated_coords(q_img, s, ext)
    clean = O.nufft2(ac, x, y, z, eps=1e-12).real.reshape(n_img, -1)
    imgs = synth.poisson_images(clean, 10.0, seed=2)
    say(f"matching end to end {precision}: M={M} refs={n_ref} images={n_img} pixels={side * side}")
    q, idx = alignment.match_orientations(0, s, ext, imgs, ac, n_ref, ref_orientations=ref, precision=precision,
                                         return_index=True)
    t = time.time()
    want, D = OA.match(ref, s, ext, imgs, ac)
    say(f"  oracle matching in {time.time() - t:.1f}s")
    ok = OA.tie_band_ok(D, rel=1e-4)
    idx = np.asarray(idx)
    say(f"  outside the tie band {ok.mean():.3f}; agreement there {(idx[ok] == want[ok]).mean():.4f}, "
        f"overall {(idx == want).mean():.4f}")
    assert ok.mean() > 0.5
    assert np.array_equal(idx[ok], want[ok])
    np.testing.assert_array_equal(q, ref[idx])
