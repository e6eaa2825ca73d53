"""GPU type-2 / type-1 NUFFT against the FP64 oracle (exact sums), the reference's
uniform-grid identity (cmtip/tests/test_nufft.py:40-96) and the adjoint identity.

Tolerance: relative L2 <= 1e-5 at eps = 1e-6 (FP32 arithmetic, kernel width 7)."""
import numpy as np
import pytest
import torch

from cmtip_amd import nufft
from oracle import nufft as O
from tests.gpu_util import dev, points, rel_l2, say

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.mark.parametrize("M", [7, 8, 21])
@pytest.mark.parametrize("cplx", [False, True])
def test_type2_vs_exact(M, cplx):
    rng = np.random.default_rng(M + 100 * cplx)
    N = 4000
    f = rng.standard_normal((M, M, M)).astype(np.float32)
    if cplx:
        f = (f + 1j * rng.standard_normal((M, M, M))).astype(np.complex64)
    x, y, z = points(rng, N)
    say(f"type2 M={M} cplx={cplx} N={N}")
    got = nufft.forward_gpu(f, x, y, z, None, use_recip_sym=not cplx)
    want = O.nudft2(f.astype(np.complex128), x, y, z)
    assert got.dtype == np.complex64 and got.shape == (N,)
    assert rel_l2(got, want) < TOL
    if not cplx:
        say(f"type2 real-out M={M}")
        re = nufft.forward_real(f, x, y, z)
        assert re.dtype == np.float32
        assert rel_l2(re, want.real) < TOL


def test_forward_support_mutates_grid():
    rng = np.random.default_rng(5)
    M = 9
    f = rng.standard_normal((M, M, M)).astype(np.float32)
    sup = (rng.random((M, M, M)) > 0.5).astype(np.float32)
    x, y, z = points(rng, 500)
    f0 = f.copy()
    say("type2 with support")
    got = nufft.forward_cpu(f, x, y, z, sup, True)
    np.testing.assert_array_equal(f, f0 * sup)  # in place, as forward.py:35
    assert rel_l2(got, O.nudft2(f0 * sup, x, y, z)) < TOL


@pytest.mark.parametrize("M", [7, 8, 21])
@pytest.mark.parametrize("kind", ["complex", "real", "ones"])
def test_type1_vs_exact(M, kind):
    rng = np.random.default_rng(M + 7)
    N = 4000
    x, y, z = points(rng, N)
    sup = (rng.random((M, M, M)) > 0.3).astype(np.float32)
    if kind == "complex":
        c = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    elif kind == "real":
        c = rng.standard_normal(N).astype(np.float32)
    else:
        c = np.ones(N, np.float32)
    want = O.nudft1(c.astype(np.complex128), x, y, z, M)
    say(f"type1 M={M} kind={kind}")
    if kind == "ones":
        got = nufft.adjoint_ones(x, y, z, M)
        assert rel_l2(got, want.real / M ** 3) < TOL
        return
    got = nufft.adjoint_gpu(c, x, y, z, M, use_recip_sym=True, support=sup)
    assert got.dtype == np.float32 and got.shape == (M, M, M)
    assert rel_l2(got, (want * sup).real / M ** 3) < TOL
    got_c = nufft.adjoint_cpu(c, x, y, z, M, use_recip_sym=False)
    assert got_c.dtype == np.complex64
    assert rel_l2(got_c, want / M ** 3) < TOL


def test_uniform_grid_identity():
    # cmtip/tests/test_nufft.py:31-49: forward on the uniform grid equals the centred DFT
    M = 21
    rng = np.random.default_rng(0)
    ls = np.linspace(-1, 1, M + 1)
    ls = (ls[:-1] + ls[1:]) / 2
    H, K, L = (np.pi * g.ravel().astype(np.float32) for g in np.meshgrid(ls, ls, ls, indexing="ij"))
    ac = np.abs(rng.standard_normal((M, M, M))).astype(np.float32)
    calc_np = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(ac)).real)
    say("uniform grid forward/adjoint")
    calc = nufft.forward_gpu(ac, H, K, L).real
    assert np.mean(np.abs(calc_np.ravel() - calc)) < 1e-4 * np.abs(calc_np).mean()
    ivol = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(ac))).real.astype(np.float32)
    back = nufft.adjoint_gpu(ivol.ravel(), H, K, L, M)
    want = np.real(np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(ivol))))
    assert rel_l2(back, want) < TOL


def test_adjoint_identity_large():
    # size-independent property at a million points: <A f, c> = <f, A^H c>
    M, N = 41, 1_000_000
    rng = np.random.default_rng(11)
    d = dev()
    f = torch.randn((M, M, M), device=d)
    c = torch.complex(torch.randn(N, device=d), torch.randn(N, device=d))
    x, y, z = (torch.empty(N, device=d).uniform_(-np.pi, np.pi) for _ in range(3))
    say("adjoint identity M=41 N=1e6")
    Af = nufft.forward_gpu(f, x, y, z)
    Ahc = nufft.adjoint_gpu(c, x, y, z, M, use_recip_sym=False) * M ** 3
    lhs = torch.vdot(Af.to(torch.complex128), c.to(torch.complex128))
    rhs = torch.vdot(f.to(torch.complex128).reshape(-1), Ahc.to(torch.complex128).reshape(-1))
    assert float(abs(lhs - rhs) / abs(lhs)) < 1e-5


@pytest.mark.parametrize("eps,tol", [(1e-4, 2e-4), (1e-5, 2e-5), (1e-6, 1e-5)])
def test_eps_controls_error(eps, tol):
    M, N = 16, 3000
    rng = np.random.default_rng(3)
    f = rng.standard_normal((M, M, M)).astype(np.float32)
    x, y, z = points(rng, N)
    say(f"type2 eps={eps}")
    got = nufft.forward_gpu(f, x, y, z, eps=eps)
    assert rel_l2(got, O.nudft2(f, x, y, z)) < tol


def test_empty_points():
    M = 8
    f = np.ones((M, M, M), np.float32)
    e = np.zeros(0, np.float32)
    say("empty point set")
    assert nufft.forward_gpu(f, e, e, e).shape == (0,)
    assert np.all(nufft.adjoint_gpu(np.zeros(0, np.complex64), e, e, e, M) == 0)


def test_out_of_range_asserts():
    # forward.py:27,64 and adjoint.py:29,66 assert max |coordinate| < 3 pi; the maximum is
    # recorded while setpts bins the points (NaN counts as out of range)
    M = 8
    f = np.ones((M, M, M), np.float32)
    rng = np.random.default_rng(3)
    ok = rng.uniform(-3.0, 3.0, 5000).astype(np.float32)
    for bad in (3 * np.pi + 0.1, -3 * np.pi - 1e-3, np.nan):
        x = ok.copy()
        x[4321] = bad
        say(f"range check with {bad}")
        with pytest.raises(AssertionError):
            nufft.forward_gpu(f, ok, ok, x)
        with pytest.raises(AssertionError):
            nufft.adjoint_gpu(np.ones(5000, np.float32), x, ok, ok, M)
    # in range: the recorded maximum is the exact max |x|
    from cmtip_amd.nufft.plan import get_plan
    nufft.forward_gpu(f, ok, ok, ok)
    assert get_plan(2, M, -1).points_max_abs() == np.abs(ok).max()


@pytest.mark.parametrize("cplx", [False, True])
def test_type2_imag_output(cplx):
    # Im(T2(f)) from the real interpolation of the fine grid's imaginary part
    from cmtip_amd._lib import CMTIP_GRID_REAL, CMTIP_PTS_IMAG
    M, N = 21, 4000
    rng = np.random.default_rng(40 + cplx)
    f = rng.standard_normal((M, M, M)).astype(np.float32)
    if cplx:
        f = (f + 1j * rng.standard_normal((M, M, M))).astype(np.complex64)
    x, y, z = points(rng, N)
    want = O.nudft2(f.astype(np.complex128), x, y, z).imag
    say(f"type2 imaginary-part output cplx={cplx}")
    p = nufft.NufftPlan(2, M, -1)
    xd, yd, zd = (torch.as_tensor(v, device=dev()) for v in (x, y, z))
    p.setpts(xd, yd, zd)
    out = torch.empty(N, dtype=torch.float32, device=dev())
    p.execute(out, torch.as_tensor(f, device=dev()), (0 if cplx else CMTIP_GRID_REAL) | CMTIP_PTS_IMAG)
    assert rel_l2(out, want) < TOL


def test_lds_read_probe_values():
    # the LDS ceiling probe reads what it wrote: thread t sums the float4 entries
    # (t + 64 it) mod 4096, entry i holding (i, i + 1, i + 2, i + 3)
    from cmtip_amd._lib import check, lib, ptr, stream
    wg, iters = 3, 100
    out = torch.empty(wg * 512, device=dev())
    say("lds read probe")
    check(lib().cmtip_lds_read_probe(wg, iters, ptr(out), stream()))
    t = np.arange(512)
    idx = (t[:, None] + 64 * np.arange(iters)[None, :]) & 4095
    want = (4 * idx + 6).astype(np.float64).sum(1)
    got = out.cpu().numpy().reshape(wg, 512)
    np.testing.assert_allclose(got, np.broadcast_to(want, (wg, 512)), rtol=1e-6)


def test_setpts_skips_only_unchanged_points():
    # a repeated setpts of the same tensors is skipped; an in-place change of the
    # coordinates, new tensors or new bin settings register the points again
    from cmtip_amd._lib import CMTIP_GRID_REAL, CMTIP_PTS_REAL
    from cmtip_amd.nufft import NufftPlan
    d = dev()
    M, N = 21, 50_000
    g = torch.Generator(device=d).manual_seed(21)
    x, y, z = (torch.empty(N, device=d).uniform_(-3.0, 3.0, generator=g) for _ in range(3))
    f = torch.randn((M, M, M), device=d, generator=g)
    flags = CMTIP_GRID_REAL | CMTIP_PTS_REAL
    say("setpts skip logic")

    def fresh(xx, yy, zz):
        p = NufftPlan(2, M, -1)
        p.setpts(xx.clone(), yy.clone(), zz.clone())
        o = torch.empty(N, device=d)
        p.execute(o, f, flags)
        return o

    p = NufftPlan(2, M, -1)
    out = torch.empty(N, device=d)
    p.setpts(x, y, z)
    p.setpts(x, y, z)  # skipped
    p.execute(out, f, flags)
    assert torch.equal(out, fresh(x, y, z))
    x.mul_(0.5)  # in place: the version counter moves
    p.setpts(x, y, z)
    p.execute(out, f, flags)
    assert torch.equal(out, fresh(x, y, z))
    p.set_bins(6, 0, 0)
    p.setpts(x, y, z)  # bins again with the new settings
    p.execute(out, f, flags)
    assert rel_l2(out, fresh(x, y, z)) < 1e-6


@pytest.mark.parametrize("real", [True, False])
def test_adjoint_flag_matches_the_other_plan_type(real):
    # a plan's adjoint on its own bins gives what the plan of the other type gives on
    # the same points (same bins, same sign): the interpolation bit for bit, the real
    # spread bit for bit (exact 64-bit fixed-point sums), the complex spread up to the
    # order of the float atomics that flush its chunks into the fine grid
    from cmtip_amd._lib import CMTIP_EXEC_ADJOINT, CMTIP_GRID_REAL, CMTIP_PTS_REAL, CmtipError
    from cmtip_amd.nufft import NufftPlan
    d = dev()
    M, N = 33, 200_000
    g = torch.Generator(device=d).manual_seed(23)
    x, y, z = (torch.empty(N, device=d).uniform_(-3.1, 3.1, generator=g) for _ in range(3))
    if real:
        flags = CMTIP_GRID_REAL | CMTIP_PTS_REAL
        c = torch.randn(N, device=d, generator=g)
        f = torch.randn((M, M, M), device=d, generator=g)
        fo, co = torch.empty((M, M, M), device=d), torch.empty(N, device=d)
    else:
        flags = 0
        c = torch.randn(N, dtype=torch.complex64, device=d, generator=g)
        f = torch.randn((M, M, M), dtype=torch.complex64, device=d, generator=g)
        fo, co = (torch.empty((M, M, M), dtype=torch.complex64, device=d),
                  torch.empty(N, dtype=torch.complex64, device=d))
    say(f"adjoint flag real={real}")
    p2, p1 = NufftPlan(2, M, -1), NufftPlan(1, M, +1)
    p2.setpts(x, y, z)
    p1.setpts(x, y, z)
    want_f = torch.empty_like(fo)
    p1.execute(c, want_f, flags, None, 1.0 / M ** 3)
    p2.execute(c, fo, flags | CMTIP_EXEC_ADJOINT, None, 1.0 / M ** 3)
    if real:
        assert torch.equal(fo, want_f)  # exact integer sums: independent of the bins
    else:
        assert rel_l2(fo, want_f) < 1e-6
    want_c = torch.empty_like(co)
    p2.execute(want_c, f, flags)
    p1.execute(co, f, flags | CMTIP_EXEC_ADJOINT)
    assert torch.equal(co, want_c)
    # the plan keeps its own direction and sign afterwards
    again = torch.empty_like(co)
    p2.execute(again, f, flags)
    assert torch.equal(again, want_c)
    # a type-2 bin size beyond the spreading subgrid is refused for the adjoint
    p2.set_bins(26, 0, 0)
    p2.setpts(x, y, z)
    with pytest.raises(CmtipError):
        p2.execute(c, fo, flags | CMTIP_EXEC_ADJOINT, None, 1.0)


def test_lds_atomic_probe_values():
    # counter c receives it + 1 from every (thread t, iteration it) with
    # (t + 64 it) mod 8192 = c; thread t reports counter t
    from cmtip_amd._lib import check, lib, ptr, stream
    wg, iters = 2, 300
    out = torch.empty(wg * 512, device=dev())
    say("lds atomic probe")
    check(lib().cmtip_lds_atomic_probe(wg, iters, ptr(out), stream()))
    acc = np.zeros(8192)
    t = np.arange(512)
    for it in range(iters):
        np.add.at(acc, (t + 64 * it) & 8191, it + 1)
    got = out.cpu().numpy().reshape(wg, 512)
    np.testing.assert_allclose(got, np.broadcast_to(acc[:512], (wg, 512)), rtol=1e-6)


@pytest.mark.parametrize("M,N,signed", [(21, 4000, False), (21, 4000, True), (81, 2_000_000, False)])
def test_setpts_strengths(M, N, signed):
    # strengths carried in the point records (setpts_strengths + CMTIP_PTS_BOUND) give the
    # spread of the same strengths gathered by index (bit for bit) and the exact sums; max
    # |coordinate| is still
    # recorded; executes that need indices or other strengths are refused until the next
    # setpts, unit strengths are allowed
    from cmtip_amd._lib import (CMTIP_EXEC_ADJOINT, CMTIP_GRID_REAL, CMTIP_PTS_BOUND, CMTIP_PTS_ONES,
                                CMTIP_PTS_REAL, CmtipError)
    from cmtip_amd.nufft import NufftPlan
    d = dev()
    rng = np.random.default_rng(M + N + signed)
    xs, ys, zs = points(rng, N, lim=np.pi)
    cs = (rng.standard_normal(N) if signed else rng.random(N) * 50).astype(np.float32)
    x, y, z, c = (torch.as_tensor(v, device=d) for v in (xs, ys, zs, cs))
    flags = CMTIP_GRID_REAL | CMTIP_PTS_REAL
    p = NufftPlan(1, M, +1)
    say(f"setpts_strengths M={M} N={N} signed={signed}")
    want = torch.empty((M, M, M), device=d)
    p.setpts(x, y, z)
    p.execute(c, want, flags, None, 1.0 / M ** 3)
    got = torch.empty_like(want)
    p.setpts_strengths(x, y, z, c)
    assert abs(p.points_max_abs() - float(np.abs(np.stack([xs, ys, zs])).max())) == 0.0
    p.execute(None, got, flags | CMTIP_PTS_BOUND, None, 1.0 / M ** 3)
    # real spreads add exact 64-bit integer sums into the fine grid: the same result bit
    # for bit whatever the order of the adds or the point records' layout
    assert torch.equal(got, want)
    again = torch.empty_like(want)
    p.execute(None, again, flags | CMTIP_PTS_BOUND, None, 1.0 / M ** 3)
    assert torch.equal(again, got)
    if N <= 4000:
        exact = O.nudft1(cs.astype(np.complex128), xs, ys, zs, M).real / M ** 3
        assert rel_l2(got, exact) < TOL
    ones_b = torch.empty_like(want)
    p.execute(None, ones_b, CMTIP_GRID_REAL | CMTIP_PTS_ONES, None, 1.0)
    with pytest.raises(CmtipError):
        p.execute(c, got, flags, None, 1.0)
    with pytest.raises(CmtipError):
        p.execute(torch.empty(N, device=d), want, flags | CMTIP_EXEC_ADJOINT)
    p.setpts(x, y, z)
    with pytest.raises(CmtipError):
        p.execute(None, got, flags | CMTIP_PTS_BOUND, None, 1.0)
    ones = torch.empty_like(want)
    p.execute(None, ones, CMTIP_GRID_REAL | CMTIP_PTS_ONES, None, 1.0)
    assert torch.equal(ones_b, ones)
    # unit strengths spread with 32-bit accumulators (weight products rounded to 2^-19):
    # the 64-bit spread of explicit ones agrees to that rounding
    real1 = torch.empty_like(want)
    p.execute(torch.ones(N, device=d), real1, flags, None, 1.0)
    assert rel_l2(ones, real1) < 2e-6
    with pytest.raises(CmtipError):
        NufftPlan(2, M, -1).setpts_strengths(x, y, z, c)


@pytest.mark.parametrize("M,n_img,side", [(21, 40, 16), (81, 1000, 64)])
def test_setpts_rotated(M, n_img, side):
    # the rotated pixels registered straight from (orientations, pixel positions) make the
    # records that gen_nonuniform_positions' coordinate arrays make: the same type-1 spreads
    # (unit, gathered and bound strengths) and type-2 values bit for bit, the same max
    # |coordinate|, and agreement with the exact sums; such a plan cannot be autotuned
    from cmtip_amd import synth
    from cmtip_amd._lib import CMTIP_GRID_REAL, CMTIP_PTS_BOUND, CMTIP_PTS_ONES, CMTIP_PTS_REAL, CmtipError
    from cmtip_amd.autocorrelation import gen_nonuniform_positions
    from cmtip_amd.nufft import NufftPlan
    d = dev()
    s_np, ext = synth.detector(side)
    s = torch.as_tensor(s_np.reshape(3, -1), dtype=torch.float64, device=d).contiguous()
    q = torch.as_tensor(synth.random_quaternions(n_img, seed=M), device=d)
    H, K, L = (t.reshape(-1) for t in gen_nonuniform_positions(q, s, recip_extent=ext))
    N = H.numel()
    rng = np.random.default_rng(M)
    c = torch.as_tensor(rng.standard_normal(N).astype(np.float32), device=d)
    say(f"setpts_rotated M={M} N={N}")
    p_arr, p_rot = NufftPlan(1, M, +1), NufftPlan(1, M, +1)
    flags = CMTIP_GRID_REAL | CMTIP_PTS_REAL
    outs = []
    for p, reg in ((p_arr, lambda cc: p_arr.setpts(H, K, L) if cc is None else p_arr.setpts_strengths(H, K, L, cc)),
                   (p_rot, lambda cc: p_rot.setpts_rotated(q, s, ext, cc))):
        got = []
        reg(None)
        for fl, cc in ((CMTIP_GRID_REAL | CMTIP_PTS_ONES, None), (flags, c)):
            f = torch.empty((M, M, M), device=d)
            p.execute(cc, f, fl, None, 1.0 / M ** 3)
            got.append(f)
        reg(c)
        f = torch.empty((M, M, M), device=d)
        p.execute(None, f, flags | CMTIP_PTS_BOUND, None, 1.0 / M ** 3)
        got.append(f)
        got.append(p.points_max_abs())
        outs.append(got)
    for a, b in zip(*outs):
        assert (a == b) if isinstance(a, float) else torch.equal(a, b)
    with pytest.raises(CmtipError):
        p_rot.autotune(c, torch.empty((M, M, M), device=d), flags)
    if N <= 20000:
        xs, ys, zs = (t.cpu().numpy() for t in (H, K, L))
        exact = O.nudft1(c.cpu().numpy().astype(np.complex128), xs, ys, zs, M).real / M ** 3
        assert rel_l2(outs[1][2], exact) < TOL
    # type 2 from the same records
    f = torch.as_tensor(rng.standard_normal((M, M, M)).astype(np.float32), device=d)
    t2 = []
    for reg in (lambda p: p.setpts(H, K, L), lambda p: p.setpts_rotated(q, s, ext)):
        p = NufftPlan(2, M, -1)
        reg(p)
        out = torch.empty(N, device=d)
        p.execute(out, f, CMTIP_GRID_REAL | CMTIP_PTS_REAL)
        t2.append(out)
    assert torch.equal(t2[0], t2[1])
    with pytest.raises(CmtipError):
        NufftPlan(2, M, -1).setpts_rotated(q, s, ext, c)


@pytest.mark.parametrize("signed", [False, True])
def test_spread_independent_of_point_order(signed):
    # type-1 spreads of real and unit strengths over orientation-major image points
    # (2.05 M points, dense bins of several chunks at M = 41) give the same grid bit for
    # bit in input order, reversed and randomly permuted: the chunks' integer sums do not
    # depend on which chunk or lane a point lands in
    from cmtip_amd import synth
    from cmtip_amd._lib import CMTIP_GRID_REAL, CMTIP_PTS_BOUND, CMTIP_PTS_ONES, CMTIP_PTS_REAL
    from cmtip_amd.autocorrelation import gen_nonuniform_positions
    from cmtip_amd.nufft import NufftPlan
    d = dev()
    M = 41
    s, ext = synth.detector(64)
    q = synth.random_quaternions(500, seed=7)
    x, y, z = gen_nonuniform_positions(torch.from_numpy(q).to(d), torch.from_numpy(s).to(d), recip_extent=ext)
    x, y, z = (t.reshape(-1).contiguous() for t in (x, y, z))
    N = x.numel()
    g = torch.Generator(device="cpu").manual_seed(11 + signed)
    c = (torch.randn(N, generator=g) if signed else torch.rand(N, generator=g) * 20).to(d)
    say(f"spread order invariance N={N} signed={signed}")
    flags = CMTIP_GRID_REAL | CMTIP_PTS_REAL
    orders = [None, torch.arange(N - 1, -1, -1, device=d), torch.randperm(N, generator=g).to(d)]
    res = []
    for perm in orders:
        xs, ys, zs, cs = (x, y, z, c) if perm is None else (t[perm].contiguous() for t in (x, y, z, c))
        p = NufftPlan(1, M, +1)
        p.setpts(xs, ys, zs)
        gathered = torch.empty((M, M, M), device=d)
        p.execute(cs, gathered, flags, None, 1.0)
        ones = torch.empty_like(gathered)
        p.execute(None, ones, CMTIP_GRID_REAL | CMTIP_PTS_ONES, None, 1.0)
        p.setpts_strengths(xs, ys, zs, cs)
        bound = torch.empty_like(gathered)
        p.execute(None, bound, flags | CMTIP_PTS_BOUND, None, 1.0)
        assert torch.equal(bound, gathered)
        res.append((gathered, ones))
    for gathered, ones in res[1:]:
        assert torch.equal(gathered, res[0][0])
        assert torch.equal(ones, res[0][1])
    assert float(res[0][0].abs().sum()) > 0.0
