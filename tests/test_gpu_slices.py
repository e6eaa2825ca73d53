"""Rotated pixel coordinates and fused slicing (compute_slices, align.py:39-65)
against the oracle's float32 coordinates (bit-exact) and FP64 slices (rel L2 <= 1e-5)."""
import numpy as np
import pytest
import torch

from cmtip_amd import alignment, autocorrelation, synth
from oracle import align as OA
from oracle import geom
from tests.gpu_util import dev, rel_l2, say

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _automatic_route_after():
    # tests here force slicing routes on the shared slicing plans; later callers get the
    # automatic choice back
    yield
    from cmtip_amd.nufft import plan as NP
    for key, p in list(NP._CACHE.items()):
        if key[-1] == "slice":
            p.set_slice_route(0)


def test_rotate_coords_bit_exact():
    s, ext = synth.detector(32)
    q = synth.random_quaternions(200, seed=4)
    say("rotate_coords 200 x 1024")
    H, K, L = autocorrelation.gen_nonuniform_positions(q, s, recip_extent=ext)
    x, y, z = geom.rotated_coords(q, s, ext)
    for got, want in ((H, x), (K, y), (L, z)):
        got = np.asarray(got).ravel()
        ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
        assert ulp.max() <= 1 and (ulp > 0).mean() < 1e-3
    Hu, Ku, Lu = autocorrelation.gen_nonuniform_positions(q, s)
    Hw, Kw, Lw = geom.rotated_vectors(q, s)
    assert np.asarray(Hu).shape == (200, 1, 1024)
    assert np.abs(np.asarray(Hu) - Hw).max() <= np.abs(Hw).max() * 2 ** -23


@pytest.mark.parametrize("M,n_side", [(15, 16), (21, 32)])
def test_compute_slices_vs_oracle(M, n_side):
    s, ext = synth.detector(n_side)
    ac = synth.gaussian_ac(M, 30, seed=M)
    q = synth.random_quaternions(40, seed=M + 1)
    say(f"compute_slices M={M} P={n_side ** 2}")
    got = alignment.compute_slices(q, s, ext, ac)
    want = OA.compute_slices(q, s, ext, ac, exact=True)
    assert got.shape == (40 * n_side * n_side,) and got.dtype == np.float32
    assert rel_l2(got, want) < 1e-5


@pytest.mark.parametrize("tile_pixels", [True, False])
def test_slicer_paths_agree(tile_pixels):
    # tiled (LDS box) and untiled orders give the same slices; scattered pixels force
    # the direct-gather fallback
    M = 41
    ac = synth.gaussian_ac(M, 100, seed=3)
    s, ext = synth.detector(64)
    q = synth.random_quaternions(30, seed=9)
    say(f"slicer tile_pixels={tile_pixels}")
    sl = alignment.Slicer(ac, s, ext, tile_pixels=tile_pixels)
    got = sl.slice(q).cpu().numpy().ravel()
    want = OA.compute_slices(q, s, ext, ac, eps=1e-12)
    assert rel_l2(got, want) < 1e-5
    rng = np.random.default_rng(0)
    srand = (rng.standard_normal((3, 1, 3000)) * ext / 3).astype(np.float32)
    srand = np.clip(srand, -ext / 2, ext / 2)
    say("slicer scattered pixels")
    sl2 = alignment.Slicer(ac, srand, ext, tile_pixels=tile_pixels)
    got2 = sl2.slice(q).cpu().numpy().ravel()
    want2 = OA.compute_slices(q, srand, ext, ac, eps=1e-12)
    assert rel_l2(got2, want2) < 1e-5


def test_pixel_order_blocks():
    s, _ = synth.detector(128)
    order = alignment.align.pixel_order(torch.from_numpy(s.reshape(3, -1)).to(dev())).cpu().numpy()
    assert sorted(order.tolist()) == list(range(128 * 128))
    blk = order[:64]
    r, c = blk // 128, blk % 128
    assert r.max() - r.min() == 7 and c.max() - c.min() == 7


def test_slice_bf16_and_moments():
    M, n_side = 21, 32
    s, ext = synth.detector(n_side)
    ac = synth.gaussian_ac(M, 30, seed=1)
    q = synth.random_quaternions(64, seed=2)
    d = dev()
    sl = alignment.Slicer(ac, s, ext)
    out = torch.empty((64, n_side * n_side), device=d)
    b16 = torch.empty((64, n_side * n_side), device=d, dtype=torch.int16)
    mom = torch.zeros(2, dtype=torch.float64, device=d)
    say("slice with bf16 + moments")
    sl.slice(q, out=out, out_bf16=b16, moments=mom)
    ref_b16 = out.to(torch.bfloat16).view(torch.int16)
    assert torch.equal(b16, ref_b16)
    o64 = out.double()
    assert abs(float(mom[0]) - float(o64.sum())) <= 1e-9 * float(o64.abs().sum())
    assert abs(float(mom[1]) - float((o64 * o64).sum())) <= 1e-9 * float((o64 * o64).sum())


def _slice_route(sl, q, route):
    d = dev()
    sl.plan.set_slice_route(route)
    n = q.shape[0]
    out = torch.empty((n, sl.P), device=d)
    b16 = torch.empty((n, sl.P), device=d, dtype=torch.int16)
    mom = torch.zeros(2, dtype=torch.float64, device=d)
    sl.slice(q, out=out, out_bf16=b16, moments=mom)
    return out, b16, mom, sl.plan.slice_route()[1]


@pytest.mark.parametrize("M,n_side,n_q,auto", [(41, 64, 40, 1), (41, 64, 3500, 2), (81, 128, 40, 1),
                                                (81, 128, 5000, 3), (151, 128, 40, 2), (151, 96, 40, 2)])
def test_slice_routes_agree(M, n_side, n_q, auto):
    # fused tiles, binned points (rotated coordinates, bin sort, LDS subgrids) and binned
    # tiles evaluate the same sums (their FP32 additions differ only in order, set by each
    # layout's row alignment).  The automatic choice: binned tiles when the probed tiles fit
    # the tile route's margin, the call has enough tiles per bin and the fine grid is wide
    # (M = 81 with 5000 orientations); fused tiles for small calls whose boxes fit LDS
    # (40 orientations at M = 41 and 81); binned points otherwise (M = 41 with 3500
    # orientations on a 90-cell grid; M = 151, where no fused box fits)
    ac = synth.gaussian_ac(M, 50, seed=5)
    s, ext = synth.detector(n_side)
    q = torch.as_tensor(synth.random_quaternions(n_q, seed=6), device=dev())
    sl = alignment.Slicer(ac, s, ext)
    say(f"slice routes M={M} side={n_side}")
    res = {r: _slice_route(sl, q, r) for r in (1, 2, 3, 0)}
    assert [res[r][3] for r in (1, 2, 3, 0)] == [1, 2, 3, auto]
    o1, m1 = res[1][0], res[1][2]
    for r in (2, 3):
        assert rel_l2(res[r][0], o1) < 1e-6
        assert torch.allclose(res[r][2], m1, rtol=1e-6, atol=0)
    assert torch.equal(res[0][0], res[auto][0])
    for o, b, _, _ in res.values():
        assert torch.equal(b, o.to(torch.bfloat16).view(torch.int16))
    # spot check of the binned routes against exact FP64 sums (coordinates of the sampled
    # orientations only)
    rng = np.random.default_rng(M)
    idx = np.sort(rng.choice(n_q * sl.P, 100, replace=False))
    qn = q.cpu().numpy()
    xs, ys, zs = [], [], []
    for o in np.unique(idx // sl.P):
        pix = idx[idx // sl.P == o] % sl.P
        xo, yo, zo = geom.rotated_coords(qn[o:o + 1], s, ext)
        xs.append(xo[pix]), ys.append(yo[pix]), zs.append(zo[pix])
    x, y, z = np.concatenate(xs), np.concatenate(ys), np.concatenate(zs)
    from oracle import nufft as ON
    want = ON.nudft2(ac.astype(np.float64), x, y, z).real
    for r in (2, 3):
        assert rel_l2(res[r][0].cpu().numpy().ravel()[idx], want) < 1e-5


def test_tile_route_overflow_tiles():
    # pixels in index order (no 8 x 8 blocking): most 64-pixel tiles are detector rows,
    # wider than the tile route's margin, and go through its direct-gather overflow bin
    M = 41
    ac = synth.gaussian_ac(M, 50, seed=9)
    s, ext = synth.detector(64)
    q = torch.as_tensor(synth.random_quaternions(24, seed=10), device=dev())
    sl = alignment.Slicer(ac, s, ext, tile_pixels=False)
    say("tile route with overflow tiles")
    o3 = _slice_route(sl, q, 3)[0]
    o1 = _slice_route(sl, q, 1)[0]
    assert rel_l2(o3, o1) < 1e-6
    want = OA.compute_slices(q.cpu().numpy(), s, ext, ac, eps=1e-12)
    assert rel_l2(o3.cpu().numpy().ravel(), want) < 1e-5


def test_binned_route_batches_and_replaces_points():
    # more than one batch of the binned route (2^27 points: 8192 orientations of 16384
    # pixels); afterwards the plan's points must be registered again before execute
    from cmtip_amd._lib import CMTIP_GRID_REAL, CMTIP_PTS_REAL, CmtipError
    M, n = 41, 8200
    d = dev()
    ac = synth.gaussian_ac(M, 50, seed=7)
    s, ext = synth.detector(128)
    q = torch.as_tensor(synth.random_quaternions(n, seed=8), device=d)
    sl = alignment.Slicer(ac, s, ext)
    say(f"binned route batches n={n}")
    o1 = _slice_route(sl, q, 1)[0]
    o2 = _slice_route(sl, q, 2)[0]
    assert rel_l2(o2, o1) < 1e-6
    del o1, o2
    p = sl.plan
    x = torch.zeros(10, device=d)
    c = torch.empty(10, device=d)
    acd = torch.as_tensor(ac, device=d)
    with pytest.raises(CmtipError):
        p.execute(c, acd, CMTIP_GRID_REAL | CMTIP_PTS_REAL)
    p.setpts(x, x, x)
    p.execute(c, acd, CMTIP_GRID_REAL | CMTIP_PTS_REAL)
    assert abs(float(c[0]) - float(acd.double().sum())) <= 1e-5 * float(acd.double().abs().sum())


def test_tile_route_small_grid():
    # a forced tile route on a fine grid smaller than its subgrid (M = 7, n = 16): the
    # staged subgrid wraps the grid more than once
    M = 7
    ac = synth.gaussian_ac(M, 5, seed=2)
    s, ext = synth.detector(16)
    q = torch.as_tensor(synth.random_quaternions(12, seed=3), device=dev())
    sl = alignment.Slicer(ac, s, ext)
    say("tile route small grid M=7")
    o3 = _slice_route(sl, q, 3)[0]
    want = OA.compute_slices(q.cpu().numpy(), s, ext, ac, eps=1e-12)
    assert rel_l2(o3.cpu().numpy().ravel(), want) < 1e-5


def test_two_slicers_of_one_size_keep_their_volumes():
    # Slicers of the same M share one cached plan; interleaved calls must each slice
    # their own autocorrelation (align.py:39-65 slices the ac it is given)
    M, side = 21, 32
    s, ext = synth.detector(side)
    acs = [synth.gaussian_ac(M, 40, seed=31), synth.gaussian_ac(M, 40, seed=32)]
    q = synth.random_quaternions(8, seed=33)
    say("two slicers sharing a plan")
    a = alignment.Slicer(acs[0], s, ext)
    b = alignment.Slicer(acs[1], s, ext)
    got_a, got_b = a.slice(q).cpu().numpy(), b.slice(q).cpu().numpy()
    got_a2 = a.slice(q).cpu().numpy()
    for ac, got in ((acs[0], got_a), (acs[1], got_b), (acs[0], got_a2)):
        want = OA.compute_slices(q, s, ext, ac, exact=True).reshape(8, -1)
        assert rel_l2(got, want) < 1e-5


@pytest.mark.parametrize("route", [1, 2, 3])
def test_slice_moments_deterministic(route):
    # the per-batch std ratio (align.py:122-124) comes from these moments: every route
    # sums them in a fixed order, so repeated calls give the same bits
    M, n_side, n_q = 81, 128, 300
    s, ext = synth.detector(n_side)
    ac = synth.gaussian_ac(M, 300, seed=5)
    q = synth.random_quaternions(n_q, seed=6)
    sl = alignment.Slicer(ac, s, ext)
    say(f"moments twice, route {route}")
    out1, _, mom1, used = _slice_route(sl, q, route)
    out2, _, mom2, _ = _slice_route(sl, q, route)
    assert used == route
    assert torch.equal(out1, out2)
    assert torch.equal(mom1, mom2)
    o64 = out1.double()
    assert abs(float(mom1[0]) - float(o64.sum())) <= 1e-12 * float(o64.abs().sum())
    assert abs(float(mom1[1]) - float((o64 * o64).sum())) <= 1e-12 * float((o64 * o64).sum())
