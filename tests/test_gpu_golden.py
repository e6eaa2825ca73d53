"""The HIP path against the committed golden fixtures (tests/golden/*.npz): exact
NUDFT sums, exact slices, the FP64 distance table's assignments, the Toeplitz and
projection operators and a 5-iteration CG solve."""
import os

import numpy as np
import pytest
import torch

from cmtip_amd import alignment, nufft
from cmtip_amd import autocorrelation as A
from cmtip_amd._lib import CMTIP_GRID_REAL, CMTIP_PTS_REAL
from tests.gpu_util import dev, rel_l2, say

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-5  # FP32 NUFFT at eps 1e-6 against exact sums


def _load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def test_nufft_vs_golden():
    g = _load("nufft_m21")
    M = int(g["M"])
    d = dev()
    x, y, z = (torch.as_tensor(g[k], device=d) for k in ("x", "y", "z"))
    say("type 2 / type 1 vs golden exact sums (M=21)")
    p2 = nufft.NufftPlan(2, M, -1)
    p2.setpts(x, y, z)
    out = torch.empty(x.numel(), dtype=torch.complex64, device=d)
    p2.execute(out, torch.as_tensor(g["f"], device=d), CMTIP_GRID_REAL)
    assert rel_l2(out, g["t2"]) < TOL
    p1 = nufft.NufftPlan(1, M, +1)
    p1.setpts(x, y, z)
    f = torch.empty((M,) * 3, dtype=torch.complex64, device=d)
    p1.execute(torch.as_tensor(g["c"], device=d), f, 0, None, 1.0)
    assert rel_l2(f, g["t1"]) < TOL
    # the real-output forms keep the real parts
    re = torch.empty(x.numel(), device=d)
    p2.execute(re, torch.as_tensor(g["f"], device=d), CMTIP_GRID_REAL | CMTIP_PTS_REAL)
    assert rel_l2(re, g["t2"].real) < TOL


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_slices_and_match_vs_golden(precision):
    g = _load("slices_match_m21")
    s, ext, ac = g["s"], float(g["ext"]), g["ac"]
    say(f"slices and matching vs golden ({precision})")
    sl = alignment.compute_slices(g["q_sl"], s, ext, ac).reshape(g["slices"].shape)
    assert rel_l2(sl, g["slices"]) < TOL
    _, idx = alignment.match_orientations(0, s, ext, g["imgs"], ac, len(g["q_ref"]), ref_orientations=g["q_ref"],
                                          precision=precision, return_index=True)
    ok = g["ok"]
    assert np.array_equal(np.asarray(idx)[ok], g["argmin"][ok])


def test_ac_operators_and_cg_vs_golden():
    g = _load("ac_m9")
    M = int(g["M"])
    say("Toeplitz, projection and 5 CG iterations vs golden (M=9)")
    got = A.core_problem_convolution(g["u"], M, g["F"], 2 * M, g["S"])
    assert rel_l2(got, g["conv"]) < TOL
    anti = A.gen_F_antisupport(M)
    got = A.fourier_reg(g["u"], g["S"], anti, M, True)
    assert rel_l2(got, g["reg"]) < TOL
    ac5, info = A.solve_ac(0, g["s"], float(g["ext"]), g["data"], M, orientations=g["q"], maxiter=5, rlambda=1e-4,
                           rtol=0.0, return_info=True)
    assert info["iterations"] == int(g["iters"]) == 5
    assert rel_l2(ac5, g["ac5"]) < 1e-4
