"""Bin / launch settings of the binned NUFFT kernels and their autotuner.

Every setting must give the same transform: results are compared with the FP64 oracle
(exact sums, relative L2 <= 1e-5 at eps = 1e-6) and with the default settings."""
import numpy as np
import pytest
import torch

from cmtip_amd import nufft
from cmtip_amd._lib import CMTIP_GRID_REAL, CMTIP_PTS_ONES, CMTIP_PTS_REAL, CmtipError
from cmtip_amd.nufft import NufftPlan, spread_candidates
from oracle import nufft as O
from tests.gpu_util import dev, points, rel_l2, say

pytestmark = pytest.mark.gpu
TOL = 1e-5

# (bin_size, chunk, wg_per_cu); 0 = default.  M = 21 has a 48^3 fine grid and w = 7.
SETTINGS = [(2, 64, 1), (4, 1024, 2), (9, 0, 0), (14, 4096, 4), (0, 128, 8), (12, 333, 3)]


def _tensors(*arrs):
    return [torch.as_tensor(a, device=dev()) for a in arrs]


@pytest.mark.parametrize("setting", SETTINGS)
@pytest.mark.parametrize("kind", ["complex", "real", "ones"])
def test_type1_settings_vs_exact(setting, kind):
    M, N = 21, 4000
    rng = np.random.default_rng(31)
    x, y, z = points(rng, N)
    if kind == "complex":
        c = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
        flags = 0
    elif kind == "real":
        c = rng.standard_normal(N).astype(np.float32)
        flags = CMTIP_PTS_REAL
    else:
        c = np.ones(N, np.float32)
        flags = CMTIP_PTS_ONES
    want = O.nudft1(c.astype(np.complex128), x, y, z, M)
    say(f"type1 settings={setting} kind={kind}")
    p = NufftPlan(1, M, +1)
    p.set_bins(*setting)
    xd, yd, zd = _tensors(x, y, z)
    p.setpts(xd, yd, zd)
    f = torch.empty((M, M, M), dtype=torch.complex64, device=dev())
    cd = None if kind == "ones" else torch.as_tensor(c, device=dev())
    p.execute(cd, f, flags, None, 1.0)
    assert rel_l2(f, want) < TOL


@pytest.mark.parametrize("setting", SETTINGS)
@pytest.mark.parametrize("cplx", [False, True])
def test_type2_settings_vs_exact(setting, cplx):
    M, N = 21, 4000
    rng = np.random.default_rng(32)
    f = rng.standard_normal((M, M, M)).astype(np.float32)
    if cplx:
        f = (f + 1j * rng.standard_normal((M, M, M))).astype(np.complex64)
    x, y, z = points(rng, N)
    want = O.nudft2(f.astype(np.complex128), x, y, z)
    say(f"type2 settings={setting} cplx={cplx}")
    p = NufftPlan(2, M, -1)
    p.set_bins(*setting)
    xd, yd, zd = _tensors(x, y, z)
    p.setpts(xd, yd, zd)
    out = torch.empty(N, dtype=torch.complex64, device=dev())
    p.execute(out, torch.as_tensor(f, device=dev()), 0 if cplx else CMTIP_GRID_REAL)
    assert rel_l2(out, want) < TOL


def test_set_bins_rejects_and_reports():
    say("set_bins argument checks")
    p1 = NufftPlan(1, 21, +1)
    assert p1.get_bins() == (10, 4096, 0)  # default: 16^3 subgrid at w = 7
    for bad in [(16, 0, 0), (49, 0, 0), (-1, 0, 0), (0, 32, 0), (0, 5000, 0), (0, 0, 17), (0, 0, -1)]:
        with pytest.raises(CmtipError):
            p1.set_bins(*bad)
    assert p1.get_bins() == (10, 4096, 0)  # rejected settings leave the plan unchanged
    p1.set_bins(14, 100, 2)  # 20^3 complex subgrid: the largest that fits beside the dealing arrays
    assert p1.get_bins() == (14, 100, 2)
    with pytest.raises(CmtipError):
        p1.set_bins(15, 0, 0)
    p2 = NufftPlan(2, 21, -1)
    p2.set_bins(26, 0, 0)  # float interpolation subgrids (32^3 cells) fit a larger bin than spreading ones
    with pytest.raises(CmtipError):
        p2.set_bins(27, 0, 0)


@pytest.mark.parametrize("nufft_type", [1, 2])
def test_autotune_keeps_result(nufft_type):
    M, N = 41, 300_000
    d = dev()
    g = torch.Generator(device=d).manual_seed(7)
    x, y, z = (torch.empty(N, device=d).uniform_(-np.pi, np.pi, generator=g) for _ in range(3))
    say(f"autotune type={nufft_type} M={M} N={N}")
    ref = NufftPlan(nufft_type, M, -1 if nufft_type == 2 else +1)
    ref.setpts(x, y, z)
    tuned = NufftPlan(nufft_type, M, -1 if nufft_type == 2 else +1)
    tuned.setpts(x, y, z)
    if nufft_type == 1:
        c = torch.randn(N, device=d, generator=g)
        flags = CMTIP_PTS_REAL | CMTIP_GRID_REAL
        f_ref = torch.empty((M, M, M), device=d)
        ref.execute(c, f_ref, flags, None, 1.0)
        f = torch.empty_like(f_ref)
        best, table = tuned.autotune(c, f, flags, None, 1.0, reps=2)
        f.zero_()
        tuned.execute(c, f, flags, None, 1.0)
        got, want = f, f_ref
    else:
        f = torch.randn((M, M, M), device=d, generator=g)
        flags = CMTIP_GRID_REAL | CMTIP_PTS_REAL
        c_ref = torch.empty(N, device=d)
        ref.execute(c_ref, f, flags)
        c = torch.empty_like(c_ref)
        best, table = tuned.autotune(c, f, flags, reps=2)
        c.zero_()
        tuned.execute(c, f, flags)
        got, want = c, c_ref
    say(f"autotune best={best}")
    assert len(table) >= len(spread_candidates(ref.w, nufft_type))
    timed = [r for r in table if r[4] >= 0]
    assert timed and all(r[3] > 0 and r[4] > 0 for r in timed)
    assert tuple(best) in [tuple(r[:3]) for r in timed]
    bs, ch, wpc = tuned.get_bins()
    assert (bs, wpc) == (best[0] or bs, best[2]) and ch == (best[1] or 4096)
    # another bin layout changes only the order of FP32 sums
    assert rel_l2(got, want) < 1e-6


def test_explicit_candidates_and_misfits():
    M, N = 21, 20_000
    d = dev()
    g = torch.Generator(device=d).manual_seed(8)
    x, y, z = (torch.empty(N, device=d).uniform_(-np.pi, np.pi, generator=g) for _ in range(3))
    p = NufftPlan(1, M, +1)
    p.setpts(x, y, z)
    f = torch.empty((M, M, M), device=d)
    say("autotune explicit candidates")
    best, table = p.autotune(None, f, CMTIP_PTS_ONES | CMTIP_GRID_REAL, None, 1.0,
                             candidates=[(40, 0, 0), (6, 512, 1), (8, 0, 2)], reps=1, setpts_weight=1.0)
    assert table[0][3] == -1 and table[0][4] == -1  # 46^3 subgrid: does not fit LDS
    assert best in [(6, 512, 1), (8, 0, 2)]
    with pytest.raises(CmtipError):
        p.autotune(None, f, CMTIP_PTS_ONES | CMTIP_GRID_REAL, None, 1.0, candidates=[(40, 0, 0)], reps=1)
    # a failed tuning restores the previous settings and binning
    assert p.get_bins()[:1] == (best[0],)
    p.execute(None, f, CMTIP_PTS_ONES | CMTIP_GRID_REAL, None, 1.0)
    want = nufft.adjoint_ones(x, y, z, M) * M ** 3
    assert rel_l2(f, want) < 1e-6


def test_solve_ac_autotuned_matches_default():
    from cmtip_amd import synth
    from cmtip_amd.alignment import compute_slices
    from cmtip_amd.autocorrelation import solve_ac
    from cmtip_amd.nufft.plan import _TUNED
    M, n_img = 21, 64
    s, ext = synth.detector(24)
    q = synth.random_quaternions(n_img, seed=9)
    ac = synth.gaussian_ac(M, 40, seed=0)
    imgs = np.abs(compute_slices(q, s, ext, ac).reshape(n_img, -1)).astype(np.float32)
    say("solve_ac autotune on/off M=21")
    a0 = solve_ac(0, s, ext, imgs, M, orientations=q, maxiter=10, rlambda=1e-4)
    n_before = len(_TUNED)
    a1 = solve_ac(0, s, ext, imgs, M, orientations=q, maxiter=10, rlambda=1e-4, autotune=True)
    assert len(_TUNED) >= n_before + 2  # the kernel spread (ones) and the rhs spread
    assert rel_l2(a1, a0) < 1e-5
    nufft.clear_plans()
