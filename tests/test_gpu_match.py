"""Orientation matching (align.py:68-266) against the FP64 oracle: sklearn distances,
per-batch std scaling, argmin.  Assignments must be identical for every image whose
best-to-second-best squared-distance gap exceeds 1e-4 of the best (the band covers
the 1e-5 FP32 slice error and the FP32 cross-term rounding)."""
import numpy as np
import pytest
import torch

from cmtip_amd import alignment, synth
from oracle import align as OA
from oracle import nufft as O
from oracle import geom
from tests.gpu_util import dev, rel_l2, say

pytestmark = pytest.mark.gpu
BAND = 1e-4


def _dataset(M=21, n_side=32, n_ref=300, n_img=64, photons=10.0, seed=0):
    s, ext = synth.detector(n_side)
    ac = synth.gaussian_ac(M, 40, seed=seed)
    ref = synth.random_quaternions(n_ref, seed=seed + 1)
    q_img = synth.random_quaternions(n_img, seed=seed + 2)
    x, y, z = geom.rotated_coords(q_img, s, ext)
    clean = O.nufft2(ac.astype(np.float64), x, y, z, eps=1e-12).real.reshape(n_img, -1)
    imgs = synth.poisson_images(clean, photons, seed=seed + 3) if photons else clean.astype(np.float32)
    return s, ext, ac, ref, q_img, imgs


def test_calc_eudist_vs_sklearn():
    from sklearn.metrics.pairwise import euclidean_distances
    rng = np.random.default_rng(0)
    A = rng.random((150, 777)).astype(np.float32)
    B = rng.random((70, 777)).astype(np.float32)
    say("calc_eudist 150x70x777")
    D = alignment.calc_eudist(A, B)
    want = euclidean_distances(A.astype(np.float64), B.astype(np.float64))
    assert D.shape == (150, 70) and D.dtype == np.float32
    assert np.abs(D - want).max() < 1e-5 * want.max()


@pytest.mark.parametrize("precision,cross", [("fp32", "lib"), ("bf16", "lib"), ("bf16", "fused")])
@pytest.mark.parametrize("nbatches,order", [(1, 0), (3, 0), (2, -2)])
def test_match_vs_oracle(precision, cross, nbatches, order):
    s, ext, ac, ref, _, imgs = _dataset()
    say(f"match precision={precision} cross={cross} nbatches={nbatches} order={order}")
    q, idx = alignment.match_orientations(0, s, ext, imgs, ac, len(ref), nbatches=nbatches, order=order,
                                         ref_orientations=ref, precision=precision, return_index=True,
                                         cross=cross)
    want, D = OA.match(ref, s, ext, imgs, ac, nbatches=nbatches, order=order)
    ok = OA.tie_band_ok(D, rel=BAND)
    assert ok.mean() > 0.5, "too few images outside the tie band for a meaningful check"
    assert np.array_equal(np.asarray(idx)[ok], want[ok])
    np.testing.assert_array_equal(q, ref[np.asarray(idx)])
    assert q.dtype == np.float32 and q.shape == (imgs.shape[0], 4)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_true_orientations_recovered(precision):
    # align.py:109-111: the true quaternions are planted in the reference set
    s, ext, ac, ref, q_img, imgs = _dataset(n_img=24, photons=0)
    np.random.seed(7)
    say(f"true-orientation recovery {precision}")
    q = alignment.match_orientations(0, s, ext, imgs, ac, len(ref), nbatches=2, ref_orientations=ref,
                                     true_orientations=q_img, precision=precision)
    np.testing.assert_array_equal(q, q_img)


def test_match_torch_inputs_stay_on_device():
    s, ext, ac, ref, _, imgs = _dataset(n_img=16)
    d = dev()
    say("match with device tensors")
    q = alignment.match_orientations(0, s, ext, torch.from_numpy(imgs).to(d), torch.from_numpy(ac).to(d), len(ref),
                                     ref_orientations=ref)
    assert isinstance(q, torch.Tensor) and q.is_cuda


def test_match_mult_assigns_conformations():
    M, n_side = 21, 32
    s, ext = synth.detector(n_side)
    acs = [synth.gaussian_ac(M, 40, seed=10), synth.gaussian_ac(M, 40, seed=20)]
    ref = synth.random_quaternions(200, seed=3)
    q_img = synth.random_quaternions(20, seed=4)
    x, y, z = geom.rotated_coords(q_img, s, ext)
    conf = np.array([0] * 10 + [1] * 10)
    imgs = np.stack([O.nufft2(acs[c].astype(np.float64), x[i * n_side ** 2:(i + 1) * n_side ** 2],
                              y[i * n_side ** 2:(i + 1) * n_side ** 2], z[i * n_side ** 2:(i + 1) * n_side ** 2],
                              eps=1e-12).real for i, c in enumerate(conf)]).astype(np.float32)
    say("match_orientations_mult 2 volumes")
    q, got_conf = alignment.match_orientations_mult(0, s, ext, imgs, acs, len(ref), ref_orientations=ref,
                                                    true_orientations=q_img)
    np.testing.assert_array_equal(got_conf, conf)
    np.testing.assert_array_equal(q, q_img)


def test_match_batch_wrapper():
    # each image batch is aligned on its own, with its own data std (align.py:249-265)
    s, ext, ac, ref, _, imgs = _dataset(n_img=50)
    say("match_orientations_batch")
    q = alignment.match_orientations_batch(0, s, ext, imgs, ac, len(ref), batch_size=20, ref_orientations=ref)
    assert q.shape == (50, 4) and q.dtype == np.float64
    for start, end in ((0, 20), (20, 50)):
        qb = alignment.match_orientations(0, s, ext, imgs[start:end], ac, len(ref), ref_orientations=ref)
        np.testing.assert_array_equal(q[start:end], qb.astype(np.float64))


def _kernel_case(n_ref, n_img, P, seed):
    d = dev()
    g = torch.Generator(device=d).manual_seed(seed)
    A = torch.rand((n_ref, P), device=d, generator=g) * 50
    B = torch.poisson(torch.rand((n_img, P), device=d, generator=g) * 20, generator=g)
    w2 = torch.full((P,), 1.0 / P ** 2, device=d)
    gamma = torch.tensor([0.37], dtype=torch.float64, device=d)
    A64, B64, w64 = A.double(), B.double(), w2.double()
    D = ((gamma ** 2) * (A64 * A64 * w64).sum(1)[:, None] - 2 * gamma * (A64 * w64) @ B64.T
         + (B64 * B64 * w64).sum(1)[None, :])
    return A, B, w2, gamma, D


def _gap_ok(D, rel):
    s = torch.sort(D, dim=0).values
    return ((s[1] - s[0]) > rel * s[0].abs()).cpu().numpy()


@pytest.mark.parametrize("n_ref,n_img,P", [(700, 300, 4104), (10000, 5000, 16384)])
def test_match_kernels_vs_fp64(n_ref, n_img, P):
    from cmtip_amd._lib import check, lib, ptr, stream
    A, B, w2, gamma, D = _kernel_case(n_ref, n_img, P, 0)
    want = torch.argmin(D, dim=0).cpu().numpy()
    L = lib()
    d = dev()
    Bw = torch.empty_like(B)
    Bw16 = torch.empty(B.shape, dtype=torch.int16, device=d)
    check(L.cmtip_match_prepare_images(ptr(B), n_img, P, ptr(w2), ptr(Bw), ptr(Bw16), stream()))
    an = torch.empty(n_ref, device=d)
    bn = torch.empty(n_img, device=d)
    check(L.cmtip_weighted_norms(ptr(A), n_ref, P, ptr(w2), ptr(an), stream()))
    check(L.cmtip_weighted_norms(ptr(B), n_img, P, ptr(w2), ptr(bn), stream()))
    say(f"fp32 argmin kernel {n_ref}x{n_img}x{P}")
    best = torch.full((n_img,), -1, dtype=torch.int64, device=d)
    check(L.cmtip_match_argmin_f32(ptr(A), n_ref, 0, ptr(Bw), n_img, P, ptr(an), ptr(bn), ptr(gamma), ptr(best),
                                   stream()))
    got = (best & 0xFFFFFFFF).cpu().numpy()
    ok = _gap_ok(D, 1e-5)
    assert ok.mean() > 0.5
    assert np.array_equal(got[ok], want[ok]), f"fp32 mismatches: {(got[ok] != want[ok]).sum()} of {ok.sum()}"
    say(f"bf16 top-8 kernel {n_ref}x{n_img}x{P}")
    A16 = A.to(torch.bfloat16).view(torch.int16)
    cand = torch.full((n_img, 8), -1, dtype=torch.int64, device=d)
    check(L.cmtip_match_topk_bf16(ptr(A16), n_ref, 0, ptr(Bw16), n_img, P, ptr(an), ptr(bn), ptr(gamma), 8,
                                  ptr(cand), stream()))
    gref = gamma.expand(n_ref).contiguous()
    best2 = torch.full((n_img,), -1, dtype=torch.int64, device=d)
    check(L.cmtip_match_rescore(ptr(A), n_ref, ptr(B), n_img, P, ptr(w2), ptr(gref), 8, ptr(cand), ptr(best2),
                                stream()))
    got2 = (best2 & 0xFFFFFFFF).cpu().numpy()
    ok2 = _gap_ok(D, 1e-4)
    assert np.array_equal(got2[ok2], want[ok2]), f"bf16 mismatches: {(got2[ok2] != want[ok2]).sum()} of {ok2.sum()}"
    # candidate lists are sorted and hold distinct references
    c = cand.cpu().numpy().view(np.uint64)
    assert np.all(np.diff(c.astype(np.float64), axis=1) >= 0)
    # the same cross term from the library GEMM, in two reference blocks, then top-k
    say(f"bf16 library cross term + top-k {n_ref}x{n_img}x{P}")
    cand3 = torch.full((n_img, 8), -1, dtype=torch.int64, device=d)
    half = n_ref // 2
    Dt = torch.empty((n_img, n_ref - half), device=d)
    for r0, r1 in ((0, half), (half, n_ref)):
        check(L.cmtip_match_cross_bf16(ptr(A16[r0:r1]), r1 - r0, ptr(Bw16), n_img, P, ptr(Dt), stream()))
        check(L.cmtip_match_topk_cross(ptr(Dt), r1 - r0, r0, n_img, ptr(an[r0:r1]), ptr(bn), ptr(gamma), 8,
                                       ptr(cand3), stream()))
    best3 = torch.full((n_img,), -1, dtype=torch.int64, device=d)
    check(L.cmtip_match_rescore(ptr(A), n_ref, ptr(B), n_img, P, ptr(w2), ptr(gref), 8, ptr(cand3), ptr(best3),
                                stream()))
    got3 = (best3 & 0xFFFFFFFF).cpu().numpy()
    assert np.array_equal(got3[ok2], want[ok2]), f"lib mismatches: {(got3[ok2] != want[ok2]).sum()} of {ok2.sum()}"
    c3 = cand3.cpu().numpy().view(np.uint64)
    assert np.all(np.diff(c3.astype(np.float64), axis=1) >= 0)
    # both BF16 forms keep the same candidates apart from near-ties of the approximate distances
    same = np.mean([len(set(a & 0xFFFFFFFF) & set(b & 0xFFFFFFFF)) / 8 for a, b in zip(c, c3)])
    assert same > 0.9


@pytest.mark.parametrize("n_ref,n_img,P", [(3840, 2560, 2048), (3800, 2500, 2051)])
def test_fp32_split_tail_deterministic(n_ref, n_img, P):
    # 600 tiles of 128 x 128 on 512 resident workgroups: the 88 tiles of the last round are
    # split along K into 5 runs whose partials are added in run order; the argmin matches
    # FP64 outside the tie band and repeated calls (counters reset by the kernel) give
    # identical keys; aligned (P % 4 == 0) and scalar (P odd) loads
    from cmtip_amd._lib import check, lib, ptr, stream
    A, B, w2, gamma, D = _kernel_case(n_ref, n_img, P, 3)
    want = torch.argmin(D, dim=0).cpu().numpy()
    L = lib()
    d = dev()
    Bw = torch.empty_like(B)
    Bw16 = torch.empty(B.shape, dtype=torch.int16, device=d)
    check(L.cmtip_match_prepare_images(ptr(B), n_img, P, ptr(w2), ptr(Bw), ptr(Bw16), stream()))
    an = torch.empty(n_ref, device=d)
    bn = torch.empty(n_img, device=d)
    check(L.cmtip_weighted_norms(ptr(A), n_ref, P, ptr(w2), ptr(an), stream()))
    check(L.cmtip_weighted_norms(ptr(B), n_img, P, ptr(w2), ptr(bn), stream()))
    say(f"fp32 split tail {n_ref}x{n_img}x{P}")
    keys = []
    for _ in range(3):
        best = torch.full((n_img,), -1, dtype=torch.int64, device=d)
        check(L.cmtip_match_argmin_f32(ptr(A), n_ref, 0, ptr(Bw), n_img, P, ptr(an), ptr(bn), ptr(gamma),
                                       ptr(best), stream()))
        keys.append(best.cpu().numpy())
    assert np.array_equal(keys[0], keys[1]) and np.array_equal(keys[0], keys[2])
    got = keys[0] & 0xFFFFFFFF
    ok = _gap_ok(D, 1e-5)
    assert ok.mean() > 0.5
    assert np.array_equal(got[ok], want[ok]), f"mismatches: {(got[ok] != want[ok]).sum()} of {ok.sum()}"


@pytest.mark.parametrize("n_ref,n_img,P,k", [(700, 300, 4096, 8), (517, 1111, 2048, 16), (10000, 5000, 16384, 16)])
def test_bf16_kernel_paths_agree(n_ref, n_img, P, k):
    # the direct-to-LDS kernel (aligned rows, P a multiple of 64) and the register-staged
    # kernel (rows one element off 16-byte alignment) accumulate the same MFMA sequence,
    # so their candidate lists are identical; both keep the same candidates as the library
    # GEMM + top-k pass apart from near-ties of the approximate distances; reference
    # offsets carry into the keys
    from cmtip_amd._lib import check, lib, ptr, stream
    d = dev()
    L = lib()
    g = torch.Generator(device=d).manual_seed(11)
    A = torch.randn((n_ref, P), device=d, generator=g)
    B = torch.randn((n_img, P), device=d, generator=g)
    A16 = A.to(torch.bfloat16).view(torch.int16)
    B16 = B.to(torch.bfloat16).view(torch.int16)
    # unaligned copies: rows start one element past a 16-byte boundary
    A16u = torch.empty(n_ref * P + 1, dtype=torch.int16, device=d)[1:].view(n_ref, P)
    B16u = torch.empty(n_img * P + 1, dtype=torch.int16, device=d)[1:].view(n_img, P)
    A16u.copy_(A16)
    B16u.copy_(B16)
    an = (A.double() ** 2).sum(1).float()
    bn = (B.double() ** 2).sum(1).float()
    off = 3

    def run(a, b):
        cand = torch.full((n_img, k), -1, dtype=torch.int64, device=d)
        check(L.cmtip_match_topk_bf16(ptr(a), n_ref, off, ptr(b), n_img, P, ptr(an), ptr(bn), None, k, ptr(cand),
                                      stream()))
        return cand.cpu().numpy().view(np.uint64)

    say(f"bf16 kernel paths {n_ref}x{n_img}x{P} k={k}")
    c_dma, c_reg = run(A16, B16), run(A16u, B16u)
    np.testing.assert_array_equal(c_dma, c_reg)
    assert np.all(np.diff(c_dma.astype(np.float64), axis=1) >= 0)
    idx = (c_dma & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert idx.min() >= off and idx.max() < n_ref + off
    assert all(len(set(r)) == k for r in idx[:200])
    say("  library GEMM + top-k")
    cand = torch.full((n_img, k), -1, dtype=torch.int64, device=d)
    Dt = torch.empty((n_img, n_ref), device=d)
    check(L.cmtip_match_cross_bf16(ptr(A16), n_ref, ptr(B16), n_img, P, ptr(Dt), stream()))
    check(L.cmtip_match_topk_cross(ptr(Dt), n_ref, off, n_img, ptr(an), ptr(bn), None, k, ptr(cand), stream()))
    c_lib = cand.cpu().numpy().view(np.uint64)
    same = np.mean([len(set(a) & set(b)) / k for a, b in zip(idx, (c_lib & np.uint64(0xFFFFFFFF)).astype(np.int64))])
    assert same > 0.9


@pytest.mark.parametrize("split,n_ref", [(180, 300), (181, 300), (1000, 2601)])
def test_topk_cross_exact(split, n_ref):
    # keys formed and merged exactly as numpy forms them in float32 (gamma = 1), over two
    # reference batches, with repeated distances (ties keep the smaller reference index);
    # batches of a multiple of 4 references on 16-byte aligned rows take the float4 form,
    # the others the scalar form
    from cmtip_amd._lib import check, lib, ptr, stream
    rng = np.random.default_rng(5)
    d = dev()
    n_img, k = 37, 8
    Dt = np.round(rng.random((n_img, n_ref)) * 64).astype(np.float32) / 64  # many equal values
    an = (rng.random(n_ref) * 2).astype(np.float32)
    bn = (rng.random(n_img) * 2).astype(np.float32)
    L = lib()
    cand = torch.full((n_img, k), -1, dtype=torch.int64, device=d)
    an_d, bn_d = torch.as_tensor(an, device=d), torch.as_tensor(bn, device=d)
    say("top-k from a cross-term block")
    for r0, r1 in ((0, split), (split, n_ref)):
        Db = torch.as_tensor(np.ascontiguousarray(Dt[:, r0:r1]), device=d)
        check(L.cmtip_match_topk_cross(ptr(Db), r1 - r0, r0, n_img, ptr(an_d[r0:r1]), ptr(bn_d), None, k,
                                       ptr(cand), stream()))
    d2 = np.maximum(an[None, :] + (bn[:, None] - np.float32(2) * Dt), np.float32(0)).astype(np.float32)
    keys = (d2.view(np.uint32).astype(np.uint64) << np.uint64(32)) | np.arange(n_ref, dtype=np.uint64)[None, :]
    want = np.sort(keys, axis=1)[:, :k]
    got = cand.cpu().numpy().view(np.uint64)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("P", [1024, 1023])
def test_row_kernels_float4_and_scalar_forms(P):
    # centring (y = gamma x - c, its bfloat16 copy, sum w2 y^2) and image preparation
    # (w2 (b - c)) give the same values from their float4 form (P a multiple of 4, aligned
    # rows) and their scalar form (rows one float off alignment)
    from cmtip_amd._lib import check, lib, ptr, stream
    d = dev()
    g = torch.Generator(device=d).manual_seed(11)
    rows = 300
    L = lib()
    X = torch.rand((rows, P), device=d, generator=g) * 100
    c = torch.rand(P, device=d, generator=g) * 50
    w2 = torch.rand(P, device=d, generator=g)
    gamma = torch.tensor([0.37], dtype=torch.float64, device=d)
    outs = []
    say(f"row kernels P={P}")
    for off in (0, 1):
        Xb = torch.zeros(rows * P + 1, device=d)
        Xb[off:off + rows * P] = X.reshape(-1)
        cb, wb = torch.zeros(P + 1, device=d), torch.zeros(P + 1, device=d)
        cb[off:off + P], wb[off:off + P] = c, w2
        Xo, co, wo = Xb[off:off + rows * P], cb[off:off + P], wb[off:off + P]
        Yb = torch.empty(rows * P + 1, device=d)
        Y16 = torch.empty(rows * P + 4, dtype=torch.int16, device=d)
        norms = torch.empty(rows, device=d)
        check(L.cmtip_center_rows(ptr(Xo), rows, P, ptr(gamma), ptr(co), ptr(wo), ptr(Yb[off:]),
                                  ptr(Y16[4 * off:]), ptr(norms), stream()))
        Bw = torch.empty(rows * P + 1, device=d)
        Bw16 = torch.empty(rows * P + 4, dtype=torch.int16, device=d)
        check(L.cmtip_match_prepare_images_centered(ptr(Xo), rows, P, ptr(wo), ptr(co), ptr(Bw[off:]),
                                                    ptr(Bw16[4 * off:]), stream()))
        outs.append((Yb[off:off + rows * P].clone(), Y16[4 * off:4 * off + rows * P].clone(), norms.clone(),
                     Bw[off:off + rows * P].clone(), Bw16[4 * off:4 * off + rows * P].clone(), co, wo))
    (y0, h0, n0, b0, bh0, c0, w0), (y1, h1, n1, b1, bh1, _, _) = outs
    assert torch.equal(y0, y1) and torch.equal(h0, h1) and torch.equal(b0, b1) and torch.equal(bh0, bh1)
    assert float(((n0 - n1).abs() / n1.abs()).max()) < 1e-6
    # against torch: w2 (b - c) exactly; bfloat16 copies rounded to nearest even; y and its
    # FP64 norm to float32 rounding
    xb = X.reshape(-1)
    want_b = (X - c0[None, :]) * w0[None, :]
    assert torch.equal(b0, want_b.reshape(-1))
    assert torch.equal(bh0, want_b.reshape(-1).to(torch.bfloat16).view(torch.int16))
    assert torch.equal(h0, y0.to(torch.bfloat16).view(torch.int16))
    g32 = float(np.float32(0.37))  # the kernels scale by gamma rounded to float32
    y64 = g32 * xb.double() - c0.double().repeat(rows)  # g x is exact in FP64
    assert bool(((y0.double() - y64).abs() <= 1e-7 * y64.abs() + 1e-12).all())
    n64 = (w0.double()[None, :] * y0.double().reshape(rows, P) ** 2).sum(1)
    assert float(((n0.double() - n64).abs() / n64).max()) < 1e-6


@pytest.mark.parametrize("P", [1024, 1023])
def test_prepare_images_norms_one_pass(P):
    # the one-pass image preparation gives the bfloat16 images and the norms of the two
    # separate calls bit for bit (P = 1023 takes the two-call form inside)
    from cmtip_amd._lib import check, lib, ptr, stream
    d = dev()
    g = torch.Generator(device=d).manual_seed(12)
    n, L = 257, lib()
    B = torch.rand((n, P), device=d, generator=g) * 30
    c = torch.rand(P, device=d, generator=g) * 10
    w2 = torch.rand(P, device=d, generator=g)
    say(f"prepare_images_norms P={P}")
    h1 = torch.empty((n, P), dtype=torch.int16, device=d)
    n1 = torch.empty(n, device=d)
    check(L.cmtip_match_prepare_images_norms(ptr(B), n, P, ptr(w2), ptr(c), None, ptr(h1), ptr(n1), stream()))
    h2 = torch.empty_like(h1)
    n2 = torch.empty_like(n1)
    check(L.cmtip_match_prepare_images_centered(ptr(B), n, P, ptr(w2), ptr(c), None, ptr(h2), stream()))
    check(L.cmtip_center_rows(ptr(B), n, P, None, ptr(c), ptr(w2), None, None, ptr(n2), stream()))
    assert torch.equal(h1, h2) and torch.equal(n1, n2)
    want = ((B.double() - c.double()) ** 2 * w2.double()).sum(1)
    assert float(((n1.double() - want).abs() / want).max()) < 1e-6


@pytest.mark.parametrize("masked", [False, True])
def test_match_gamma_exact(masked):
    # std(images) / std(model) from (sum, sum of squares[, count]) in FP64 with the
    # reference's order of operations, rounded to float32, written to every slot
    import struct
    from cmtip_amd._lib import check, lib, ptr, stream
    d = dev()
    rng = np.random.default_rng(13 + masked)
    img = rng.random(3) * np.array([1e6, 1e9, 1.0])
    img_count = 0.0 if masked else 4.2e7
    if masked:
        img[2] = 3.1e7
    mdl = rng.random(2) * np.array([2e5, 3e7])
    model_count = 1.6384e8
    img_d = torch.tensor(img, dtype=torch.float64, device=d)
    mdl_d = torch.tensor(mdl, dtype=torch.float64, device=d)
    out = torch.zeros(37, dtype=torch.float64, device=d)
    say(f"match_gamma masked={masked}")
    check(lib().cmtip_match_gamma(ptr(img_d), img_count, ptr(mdl_d), model_count, 37, ptr(out), stream()))
    ci = max(img[2], 1.0) if masked else img_count
    mi = img[0] / ci
    si = np.sqrt(max(img[1] / ci - mi * mi, 0.0))
    mm = mdl[0] / model_count
    sm = np.sqrt(max(mdl[1] / model_count - mm * mm, 0.0))
    want = struct.unpack("f", struct.pack("f", si / sm))[0]
    assert torch.equal(out, torch.full_like(out, want))


def test_weights2_cache():
    # squared pixel weights: cached per host detector array and order (one device
    # tensor reused), recomputed for device inputs, equal to generate_weights squared
    from cmtip_amd.alignment import align as AL
    s, _ = synth.detector(32)
    say("squared weights cache")
    for order in (0, -2):
        a = AL._weights2(s, order)
        assert AL._weights2(s.copy(), order) is a
        w = AL.generate_weights(torch.as_tensor(s, device=dev()), order=order).reshape(-1)
        assert torch.equal(a, w * w)
        t = torch.as_tensor(s, device=dev())
        assert AL._weights2(t, order) is not AL._weights2(t, order)
    assert AL._weights2(s, 0) is not AL._weights2(s, -2)


def test_cross_bf16_vs_fp32_product():
    from cmtip_amd._lib import check, lib, ptr, stream
    d = dev()
    g = torch.Generator(device=d).manual_seed(3)
    n_ref, n_img, K = 1000, 600, 2 * 1234
    A16 = torch.randn(n_ref, K, device=d, generator=g).to(torch.bfloat16)
    B16 = torch.randn(n_img, K, device=d, generator=g).to(torch.bfloat16)
    Dt = torch.empty((n_img, n_ref), device=d)
    say("library bf16 cross term")
    check(lib().cmtip_match_cross_bf16(ptr(A16.view(torch.int16)), n_ref, ptr(B16.view(torch.int16)), n_img, K,
                                       ptr(Dt), stream()))
    want = B16.double() @ A16.double().T  # exact products of the bf16 values, FP64 sums
    err = (Dt.double() - want).abs().max() / want.abs().max()
    assert float(err) < 1e-5


def test_generate_weights():
    s, _ = synth.detector(16)
    for order in (0, -2, 1):
        w = alignment.generate_weights(s, order)
        want = OA.generate_weights(s, order)
        assert w.shape == (1, 256)
        assert rel_l2(w, want) < 1e-6


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_realistic_dynamic_range_vs_fp64(precision):
    # images with the benchmark's statistics (M=81 autocorrelation, 128x128 pixels,
    # Poisson at 10 photons/pixel: central pixels far brighter than the edges)
    d = dev()
    M, side, n_ref, n_img = 81, 128, 2000, 256
    s, ext = synth.detector(side)
    ac = torch.from_numpy(synth.gaussian_ac(M, 1000, seed=0)).to(d)
    ref = synth.random_quaternions(n_ref, seed=0)
    q_img = synth.random_quaternions(n_img, seed=1)
    sl = alignment.Slicer(ac, s, ext)
    clean = sl.slice(q_img)
    torch.manual_seed(2)
    imgs = torch.poisson(clean.clamp(min=0) * (10.0 / float(clean.clamp(min=0).mean())))
    say(f"realistic matching {precision}")
    q, idx = alignment.match_orientations(0, s, ext, imgs, ac, n_ref, ref_orientations=ref, precision=precision,
                                         return_index=True)
    A = sl.slice(ref).double()
    B = imgs.double()
    w2 = 1.0 / (side * side) ** 2
    gamma = float(B.std(unbiased=False) / A.std(unbiased=False))
    D = w2 * ((gamma * A) ** 2).sum(1)[:, None] - 2 * w2 * gamma * A @ B.T + w2 * (B ** 2).sum(1)[None, :]
    want = torch.argmin(D, dim=0).cpu().numpy()
    ok = _gap_ok(D, 1e-5)
    got = idx.cpu().numpy()
    assert ok.mean() > 0.9
    assert np.array_equal(got[ok], want[ok]), f"{(got[ok] != want[ok]).sum()} mismatches of {ok.sum()}"


def _masked_oracle(ref, s, ext, imgs, ac, masks, nbatches):
    """d2 = sum_p m w^2 (gamma a - b)^2 in FP64, gamma = masked data std / batch model std."""
    P = s.shape[-1]
    w2 = OA.generate_weights(s, 0).reshape(-1) ** 2
    data = imgs.astype(np.float64)
    mk = masks.astype(bool)
    std_data = data[mk].std()
    quot = len(ref) // nbatches
    D = np.zeros((len(ref), len(imgs)))
    for nb in range(nbatches):
        a = OA.compute_slices(ref[nb * quot:(nb + 1) * quot], s, ext, ac, eps=1e-12).reshape(quot, P)
        g = std_data / a.std()
        diff = g * a[:, None, :] - data[None, :, :]
        D[nb * quot:(nb + 1) * quot] = (mk[None] * w2 * diff * diff).sum(-1)
    return D


@pytest.mark.parametrize("precision,cross", [("fp32", "lib"), ("bf16", "lib"), ("bf16", "fused")])
@pytest.mark.parametrize("side", [32, 24])
def test_masked_matching_vs_oracle(precision, cross, side):
    # two detector binnings with random beamstop masks of uneven size per image
    s, ext, ac, ref, _, imgs = _dataset(n_side=side, n_ref=200, n_img=40, seed=side)
    rng = np.random.default_rng(side)
    P = side * side
    yy, xx = np.mgrid[:side, :side]
    masks = np.ones((40, P), np.uint8)
    for i in range(40):
        cy, cx = rng.integers(side // 4, 3 * side // 4, 2)
        r = rng.uniform(1, side / 4)
        masks[i] = (((yy - cy) ** 2 + (xx - cx) ** 2) > r * r).ravel()
        masks[i, rng.random(P) < 0.05] = 0
    say(f"masked matching {precision} side={side}")
    q, idx = alignment.match_orientations(0, s, ext, imgs, ac, len(ref), nbatches=2, ref_orientations=ref,
                                         precision=precision, return_index=True, masks=masks, cross=cross)
    D = _masked_oracle(ref, s, ext, imgs, ac, masks, 2)
    want = np.argmin(D, axis=0)
    srt = np.sort(D, axis=0)
    ok = (srt[1] - srt[0]) > 1e-4 * srt[0]
    assert ok.mean() > 0.5
    assert np.array_equal(np.asarray(idx)[ok], want[ok])
